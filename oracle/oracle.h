/* FP64 CPU oracle (test infrastructure only; see oracle.c). */
#ifndef ORIGAMI_ORACLE_H
#define ORIGAMI_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

orc_sys *orc_create(int n, const double *q, const double *rmin, const double *eps,
                    const double *rmin14, const double *eps14, const double *rho,
                    const double *scale, const double *mass, const double *sasa_r,
                    const double *sasa_g, int n_excl, const int *excl, int n14, const int *p14,
                    int nb, const int *bonds, const double *bond_p, int na, const int *angles,
                    const double *angle_p, int nd, const int *dih, const double *dih_p, int ni,
                    const int *imp, const double *imp_p, int nc, const int *cons,
                    const double *cons_d);
void orc_destroy(orc_sys *s);
int orc_n_atoms(const orc_sys *s);

void orc_nonbonded_forces(const orc_sys *s, const double *pos, double cutoff, double *f);
void orc_nonbonded_energy(const orc_sys *s, const double *pos, double cutoff, double *out);
double orc_pairwise_descreening(double r, double rho_i_tilde, double s_rho_j_tilde);
int orc_born_radii(const orc_sys *s, const double *pos, double cutoff, double *integral,
                   double *radii);
void orc_gb_forces(const orc_sys *s, const double *pos, const double *radii, double cutoff,
                   double *f);
void orc_gb_energy(const orc_sys *s, const double *pos, const double *radii, double *out);
void orc_bonded_forces(const orc_sys *s, const double *pos, double *f);
void orc_bonded_energy(const orc_sys *s, const double *pos, double *out);
void orc_total_force(const orc_sys *s, const double *pos, double cut_nb, double cut_gb,
                     double cut_descreen, double *f, double *radii_out, double *terms);
void orc_total_energy(const orc_sys *s, const double *pos, double cut_nb, double cut_descreen,
                      double *out);
int orc_shake(const orc_sys *s, double *pos, const double *ref, double tol_sq, int max_iters);
int orc_baoab_step(const orc_sys *s, double *pos, double *vel, double *forces, const double *xi,
                   double dt, double temperature, double gamma_ps, int use_shake,
                   double cut_nb, double cut_gb, double cut_descreen);
double orc_kinetic_energy(const orc_sys *s, const double *vel);
void orc_xoshiro_u64(uint64_t seed, int count, uint64_t *out);
void orc_splitmix64(uint64_t seed, int count, uint64_t *out);
void orc_xoshiro_from_state(uint64_t *st, int count, uint64_t *out);
void orc_philox4x32(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2,
                    uint32_t c3, uint32_t *out);
void orc_philox_normals(uint32_t k0, uint32_t k1, uint32_t sys, int n_atoms, uint32_t step,
                        double *out);

void orc_tunnel_force(int n, const double *pos, const double *origin, const double *dir, double radius,
                      double length, double k, double *f);

/* PowerSasa (oracle_sasa.c) */
void orc_psa_forces(int n, const double *pos, const double *radii, const double *gamma,
                    double *f, double *stats);
int orc_psa_atom_gradient(int n, const double *pos, const double *radii, int atom, int n_nb,
                          const int *nb, double *grad);
double orc_psa_area(int n, const double *pos, const double *radii, double *per_atom);
double orc_psa_area_from_caps(int nc, const double *axes, const double *cos_a, double radius);

/* Shrake-Rupley (oracle_sasa.c) */
void orc_sr_dots(int n, double *out);
double orc_sr_energy(int n, const double *pos, const double *radii, const double *gamma, int n_dots,
                     double *per_atom);

#ifdef __cplusplus
}
#endif
#endif
