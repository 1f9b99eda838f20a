// Learned Born-radius surrogate — host interface.
//
// Per atom: a neighbour-shell descriptor (5 element classes x 32 radial bins
// of neighbour counts over 0-12 Å, plus the intrinsic radius rho; 161
// features, zero-padded to 192) -> MLP 192 -> 128 -> 128 -> 1 (SiLU), BF16
// operands on MFMA with FP32 accumulation for the two dense layers, FP32 for
// the output layer.  Output y maps to a radius R = OG_SUR_R0 + OG_SUR_RS * y.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#define OG_SUR_F 161          // descriptor features
#define OG_SUR_K1 192         // padded input width (6 x 32)
#define OG_SUR_H 128          // hidden width
#define OG_SUR_OFF_W1 0                                   // [K1][H] input-major
#define OG_SUR_OFF_B1 (OG_SUR_OFF_W1 + OG_SUR_K1 * OG_SUR_H)
#define OG_SUR_OFF_W2 (OG_SUR_OFF_B1 + OG_SUR_H)          // [H][H] input-major
#define OG_SUR_OFF_B2 (OG_SUR_OFF_W2 + OG_SUR_H * OG_SUR_H)
#define OG_SUR_OFF_W3 (OG_SUR_OFF_B2 + OG_SUR_H)
#define OG_SUR_OFF_B3 (OG_SUR_OFF_W3 + OG_SUR_H)
#define OG_SUR_P (OG_SUR_OFF_B3 + 1)                      // 41345 parameters
#define OG_SUR_R0 1.5
#define OG_SUR_RS 0.25
#define OG_SUR_BIN 0.375      // Å, 32 bins up to 12 Å
#define OG_SUR_BMAX 8192      // largest training batch

struct og_sur {
    int n_atoms = 0, n_pad = 0, n_tiles = 0, n_sys = 0;
    hipStream_t stream = nullptr;
    // geometry / inputs
    const double* pos = nullptr;       // engine-owned
    const double* rho = nullptr;       // engine-owned
    const int* tile_atom0 = nullptr;   // engine-owned
    const int* tile_cnt = nullptr;     // engine-owned
    int* cls = nullptr;                // [N] element class 0..4
    int* tile_sys = nullptr;           // [T]
    int* sys_tile_b = nullptr;         // [S]
    int* sys_tile_e = nullptr;         // [S]
    double4* tile_bound = nullptr;     // [T]
    unsigned short* X = nullptr;       // [n_pad][K1] bf16 descriptors
    double* r_sur = nullptr;           // [N] surrogate radii
    int* mode = nullptr;               // [S] 1 = system uses surrogate radii
    unsigned* dev = nullptr;           // [S] max relative deviation (float bits), last check
    // parameters (FP32 master) and BF16 copies for the MFMA layers
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    unsigned short *W1t = nullptr;     // [H][K1] output-major
    unsigned short *W2t = nullptr;     // [H][H] output-major
    unsigned short *W2i = nullptr;     // [H][H] input-major
    int adam_t = 0;
    // training set
    long long cap = 0, n_train = 0;
    unsigned short* Xtr = nullptr;     // [cap][K1]
    float* ttr = nullptr;              // [cap] normalised targets (R - R0) / RS
    // batch scratch
    unsigned short *Xb = nullptr, *XbT = nullptr, *H1T = nullptr, *dZ1T = nullptr, *dZ2T = nullptr;
    float *dZ1Tf = nullptr, *dZ2Tf = nullptr, *H2Tf = nullptr, *tb = nullptr, *dy = nullptr;
    float *lossr = nullptr, *loss1 = nullptr;
    // fallback checking
    bool enabled = false;
    int check_every = 0, calls = 0;
    double tol = 0.02;
};

// mass -> class (H, C, N, O, S/P) for the descriptor
int og_sur_create(og_sur* s, int n_atoms, int n_sys, const std::vector<double>& mass,
                  const std::vector<int>& sys_tile_off, const double* pos_dev, const double* rho_dev,
                  const int* tile_atom0_dev, const int* tile_cnt_dev, int n_tiles, long long capacity,
                  unsigned long long seed, hipStream_t stream);
void og_sur_free(og_sur* s);
const char* og_sur_error();
// descriptors of every atom (all = true) or of surrogate-mode systems only
int og_sur_descriptors(og_sur* s, bool all);
// MLP forward on the descriptors -> r_sur
int og_sur_infer(og_sur* s);
// append (descriptor, plain radius) rows for every atom; radii_dev = plain radii
int og_sur_collect(og_sur* s, const double* radii_dev);
// one Adam step on a sampled batch; loss (mean squared, normalised units) if not null
int og_sur_train_step(og_sur* s, int batch, double lr, unsigned long long seed, float* loss);
// gradient of the mean squared loss on an explicit batch (x: n x 161, target radii)
int og_sur_gradient(og_sur* s, const float* x, const float* r_target, int n, float* grad, double* loss);
// zero the per-system deviations / turn systems above tol back to plain GB
int og_sur_check_begin(og_sur* s);
int og_sur_check_apply(og_sur* s);
int og_sur_set_params(og_sur* s, const float* p);
int og_sur_get_params(og_sur* s, float* p);
