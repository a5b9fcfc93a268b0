// The learned coarse-grid correction network (no counterpart in the
// reference): a bias-free 3-level U-Net of periodic 3x3x3 convolutions,
// channels-last BF16 activations and weights, FP32 accumulation on the
// gfx950 matrix cores (v_mfma_f32_16x16x16_bf16), FP32 master weights and
// Adam.  Layers, parameter order and loss are those of the CPU oracle's
// restatement (L0..L10: cin->c1, c1->c1, pool, c1->c2, c2->c2, pool,
// c2->c3, c3->c3, up+cat, c3+c2->c2, c2->c2, up+cat, c2+c1->c1, c1->c1,
// c1->cout; W[co][tap][ci], tap = (dz+1)*9 + (dy+1)*3 + (dx+1)).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "lmm_common.h"

namespace lmm {

struct CnnConfig {
  int cin = 8, c1 = 16, c2 = 32, c3 = 64, cout = 3;
};

// Output and target tensors carry 4 channels (cout <= 4; unused ones zero).
constexpr int CNN_OUT_STRIDE = 4;

struct CnnNet;

CnnNet *cnn_new(const CnnConfig &cfg, uint64_t seed, hipStream_t s, std::string *err);
void cnn_free(CnnNet *net);
long cnn_nparams(const CnnNet *net);
const CnnConfig &cnn_config(const CnnNet *net);
int cnn_get_params(CnnNet *net, float *host);
int cnn_set_params(CnnNet *net, const float *host);
// grid n = (nx, ny, nz) of the input, each a multiple of 4; nb batch items;
// x: BF16 [nb][nz][ny][nx][cin] on the device; y: FP32 [..][4] on the device
int cnn_forward(CnnNet *net, int nb, const int n[3], const uint16_t *x, float *y);
// forward, loss sum((y - t)^2)/(2*nb*voxels*cout) into *loss_dev (device
// double), gradient into the network's FP32 gradient buffer
int cnn_loss_grad(CnnNet *net, int nb, const int n[3], const uint16_t *x, const float *t, double *loss_dev);
float *cnn_grad_dev(CnnNet *net);
int cnn_adam_step(CnnNet *net, double lr, int step);
// training pairs held on the device: inputs BF16 [cin] and targets FP32 [4]
// per voxel of grid n
int cnn_dataset_reserve(CnnNet *net, long count, const int n[3]);
long cnn_dataset_count(const CnnNet *net);
long cnn_dataset_capacity(const CnnNet *net);
const int *cnn_dataset_grid(const CnnNet *net);
void cnn_dataset_push(CnnNet *net);  // the slot cnn_dataset_slot() returned is now filled
uint16_t *cnn_dataset_x(CnnNet *net, long i);
float *cnn_dataset_t(CnnNet *net, long i);
// steps of Adam on random batches (with replacement) of the dataset; loss of
// each step into loss_host (may be null); pairs [held0, count) are left out
int cnn_train(CnnNet *net, int steps, int batch, double lr, uint64_t seed, long held0, float *loss_host);
// mean over pairs [i0, i1) of |t - y| / |t| and of |t| / |t| (= 1): the
// relative error of the network's prediction
int cnn_eval(CnnNet *net, long i0, long i1, double *rel_err);
// host tensors (FP32 in, rounded to BF16 on upload): x [nb][nz][ny][nx][cin],
// y / t [..][cout]; the gradient in the flat parameter layout
int cnn_forward_host(CnnNet *net, int nb, const int n[3], const float *x, float *y);
int cnn_loss_grad_host(CnnNet *net, int nb, const int n[3], const float *x, const float *t, double *loss, float *grad);
hipStream_t cnn_stream(CnnNet *net);
void cnn_set_stream(CnnNet *net, hipStream_t s);

// ---- features of the Stokes coarse-grid correction (3D, FP32) ----
// The residual r = (r_u, r_p) of a preconditioned Stokes vector on one
// multigrid level, its scale in velocity units s = sqrt(mean(r_u^2) dx^4 +
// mean(r_p^2) dx^2), the network input [r_u dx^2/s, r_p dx/s, log beta,
// log theta*alpha_d] (cin = 8, BF16) and target [e_u/s, e_p dx/s] (FP32, 4
// channels); the prediction back in the level's units with the mean of its
// pressure removed.
// sums of squares of r_u and r_p into out[0], out[1] (device doubles)
void launch_stokes_ss(hipStream_t s, const Grid &g, const float *const ru[3], const float *rp, double *partial,
                      double *out);
void launch_cnn_pack(hipStream_t s, const Grid &g, const float *const ru[3], const float *rp, const float *beta,
                     const float *const alpha[3], const double *ss, uint16_t *x);
void launch_cnn_pack_target(hipStream_t s, const Grid &g, const float *const eu[3], const float *ep, const double *ss,
                            float *t);
// out = a - b (n values)
void launch_sub_f32(hipStream_t s, float *out, const float *a, const float *b, long n);
// partial: RED_BLOCKS doubles of scratch; psum: one device double
void launch_cnn_unpack(hipStream_t s, const Grid &g, const float *y, const double *ss, double *partial, double *psum,
                       float *const phu[3], float *php);

}  // namespace lmm
