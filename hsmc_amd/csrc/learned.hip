// learned.hip -- learned Widom proposal (no counterpart in the reference,
// whose insertions are uniform: compute_widom_chem_pot.c:44-80).
//
// The box is divided into M^3 voxels of side ~0.25 sigma.  Each voxel gets a
// descriptor -- the sorted distances from its centre to its 12 nearest sphere
// centres (FP32) -- and a score f_v from a 12->64->64->1 ReLU network whose
// weights are BF16 and whose products run on the MFMA units with FP32
// accumulation.  Insertions are drawn from q_v = mix * softmax(f)_v +
// (1 - mix)/V by systematic sampling (insertion i falls in the voxel whose
// cumulative-mass interval holds (i + u0)/n: voxel v receives floor(n q_v)
// or one more, with mean exactly n q_v) and placed uniformly in their voxel;
// the overlap test is the exact FP64 one.  The estimate
// P = (1/(V n)) sum_hits 1/q_v of the insertion probability is unbiased for
// any q with q_v > 0, so plain uniform insertion stays the reference.
//
// The network is trained on the GPU by the kernels below (forward, backward,
// Adam) on uniform insertions: input the descriptor of the voxel holding the
// point, target -beta * max(0, 1 - d_min) with d_min the point's distance to
// the nearest centre.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ctx.h"
#include "device_common.h"
#include "hsmc_gpu.h"

using namespace hsmc;

namespace {

constexpr int NIN = HSMC_MLP_IN;       // 12 descriptor entries
constexpr int KIN = 16;                // padded to the MFMA K
constexpr int NH = HSMC_MLP_HIDDEN;    // 64
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH,
              OW3 = OB2 + NH, OB3 = OW3 + NH, NPARAM = OB3 + 1;
static_assert(NPARAM == HSMC_MLP_NPARAM, "parameter layout");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// Network as the kernels read it: BF16 weights for the MFMA layers.
struct MlpDev {
  uint16_t W1[KIN * NH];  // [k][n], rows NIN..KIN-1 zero
  uint16_t W2[NH * NH];   // [in][out]
  float b1[NH], b2[NH], w3[NH], b3, pad[3];
};

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) {  // 4 consecutive bf16
  bf16x4 v;
  v[0] = (short)p[0];
  v[1] = (short)p[1];
  v[2] = (short)p[2];
  v[3] = (short)p[3];
  return v;
}
__device__ __forceinline__ bf16x4 ld4s(const uint16_t *p, int stride) {  // strided
  bf16x4 v;
  v[0] = (short)p[0];
  v[1] = (short)p[stride];
  v[2] = (short)p[2 * stride];
  v[3] = (short)p[3 * stride];
  return v;
}

// Copy the network into LDS (whole block).
__device__ void load_net(const MlpDev *__restrict__ net, MlpDev *s) {
  const uint32_t *src = reinterpret_cast<const uint32_t *>(net);
  uint32_t *dst = reinterpret_cast<uint32_t *>(s);
  for (int i = threadIdx.x; i < (int)(sizeof(MlpDev) / 4); i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

// Forward pass of one wave's 64 rows: X [64][KIN] bf16 in LDS -> H1 [64][NH]
// bf16 in LDS -> layer 2 on MFMA -> y[64] (ys, LDS).  When H2 != nullptr the
// post-ReLU layer-2 activations are kept (FP32) for the backward pass.
// MFMA 16x16x16 operand layout: lane l feeds A row (l%16), k 4*(l/16)..+3;
// B column (l%16), k 4*(l/16)..+3; receives D column (l%16), rows
// 4*(l/16)..+3.
__device__ void mlp_forward_wave(const MlpDev *s, const uint16_t *X, uint16_t *H1, float *H2,
                                 float *ys) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, g4 = lane >> 4;
  // layer 1: [64 x 16] x [16 x 64]
  for (int R = 0; R < 4; ++R) {
    const bf16x4 a = ld4(X + (R * 16 + l16) * KIN + 4 * g4);
    for (int C = 0; C < 4; ++C) {
      const bf16x4 b = ld4s(s->W1 + (4 * g4) * NH + C * 16 + l16, NH);
      f32x4 d = {0.f, 0.f, 0.f, 0.f};
      d = mfma16(a, b, d);
      const float bias = s->b1[C * 16 + l16];
      for (int i = 0; i < 4; ++i) {
        const float h = fmaxf(d[i] + bias, 0.f);
        H1[(R * 16 + 4 * g4 + i) * NH + C * 16 + l16] = f2bf(h);
      }
    }
  }
  wave_sync();
  // layer 2 [64 x 64] x [64 x 64], then layer 3 (64 -> 1) on the VALU
  for (int R = 0; R < 4; ++R) {
    float part[4] = {0.f, 0.f, 0.f, 0.f};
    for (int C = 0; C < 4; ++C) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int kt = 0; kt < 4; ++kt) {
        const bf16x4 a = ld4(H1 + (R * 16 + l16) * NH + kt * 16 + 4 * g4);
        const bf16x4 b = ld4s(s->W2 + (kt * 16 + 4 * g4) * NH + C * 16 + l16, NH);
        acc = mfma16(a, b, acc);
      }
      const float bias = s->b2[C * 16 + l16], w = s->w3[C * 16 + l16];
      for (int i = 0; i < 4; ++i) {
        const float h = fmaxf(acc[i] + bias, 0.f);
        if (H2) H2[(R * 16 + 4 * g4 + i) * NH + C * 16 + l16] = h;
        part[i] += h * w;
      }
    }
    for (int i = 0; i < 4; ++i) {
      float v = part[i];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 4, 64);
      v += __shfl_xor(v, 8, 64);
      if (l16 == 0) ys[R * 16 + 4 * g4 + i] = v + s->b3;
    }
  }
  wave_sync();
}

// ---- forward / backward over a batch of training rows ----------------------
// One block = one wave = 64 rows.  Gradients are accumulated into grad
// (FP32, packed like the parameters) with atomics; loss += sum (y - t)^2.
__global__ void __launch_bounds__(64)
k_mlp_train_step(const MlpDev *__restrict__ net, const float *__restrict__ Xg,
                 const float *__restrict__ Tg, int n, float inv_batch, float *__restrict__ grad,
                 float *__restrict__ loss, float *__restrict__ yout) {
  __shared__ MlpDev s;
  __shared__ uint16_t X[64 * KIN];
  __shared__ uint16_t H1[64 * NH];
  __shared__ float H2[64 * NH];  // reused for dH1 (bf16) after dH2 is formed
  __shared__ uint16_t dH2[64 * NH];
  __shared__ float ys[64], dys[64];
  load_net(net, &s);
  const int lane = threadIdx.x, l16 = lane & 15, g4 = lane >> 4;
  const int row0 = blockIdx.x * 64;
  const int row = row0 + lane;
  const bool valid = row < n;
  for (int k = 0; k < KIN; ++k)
    X[lane * KIN + k] = (valid && k < NIN) ? f2bf(Xg[(size_t)row * NIN + k]) : (uint16_t)0;
  wave_sync();
  mlp_forward_wave(&s, X, H1, H2, ys);
  // loss and dy = 2 (y - t) / B
  {
    const float y = ys[lane];
    const float t = valid ? Tg[row] : y;
    const float e = valid ? (y - t) : 0.f;
    if (valid && yout) yout[row] = y;
    dys[lane] = 2.f * e * inv_batch;
    float sq = e * e;
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
    if (lane == 0) atomicAdd(loss, sq);
  }
  wave_sync();
  // dW3[j] = sum_b H2[b][j] dy[b]; db3 = sum dy; dH2 = dy w3 (H2 > 0); db2
  {
    const int j = lane;
    float g3 = 0.f, gb2 = 0.f;
    const float w = s.w3[j];
    for (int b = 0; b < 64; ++b) {
      const float h = H2[b * NH + j];
      const float dy = dys[b];
      g3 += h * dy;
      const float d = h > 0.f ? dy * w : 0.f;
      gb2 += d;
      dH2[b * NH + j] = f2bf(d);
    }
    atomicAdd(&grad[OW3 + j], g3);
    atomicAdd(&grad[OB2 + j], gb2);
    float sdy = dys[lane];
    for (int o = 32; o > 0; o >>= 1) sdy += __shfl_xor(sdy, o, 64);
    if (lane == 0) atomicAdd(&grad[OB3], sdy);
  }
  wave_sync();
  // dW2 = H1^T dH2 (K = the 64 rows)
  for (int Ri = 0; Ri < 4; ++Ri)
    for (int Co = 0; Co < 4; ++Co) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int kt = 0; kt < 4; ++kt) {
        const bf16x4 a = ld4s(H1 + (kt * 16 + 4 * g4) * NH + Ri * 16 + l16, NH);
        const bf16x4 b = ld4s(dH2 + (kt * 16 + 4 * g4) * NH + Co * 16 + l16, NH);
        acc = mfma16(a, b, acc);
      }
      for (int i = 0; i < 4; ++i)
        atomicAdd(&grad[OW2 + (Ri * 16 + 4 * g4 + i) * NH + Co * 16 + l16], acc[i]);
    }
  // dH1 = (dH2 W2^T) (H1 > 0), stored bf16 over the H2 buffer; db1
  uint16_t *dH1 = reinterpret_cast<uint16_t *>(H2);
  wave_sync();
  for (int C = 0; C < 4; ++C) {
    float colsum = 0.f;
    for (int R = 0; R < 4; ++R) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int kt = 0; kt < 4; ++kt) {
        const bf16x4 a = ld4(dH2 + (R * 16 + l16) * NH + kt * 16 + 4 * g4);
        const bf16x4 b = ld4(s.W2 + (C * 16 + l16) * NH + kt * 16 + 4 * g4);
        acc = mfma16(a, b, acc);
      }
      for (int i = 0; i < 4; ++i) {
        const int b = R * 16 + 4 * g4 + i, c = C * 16 + l16;
        const float d = bf2f(H1[b * NH + c]) > 0.f ? acc[i] : 0.f;
        colsum += d;
        dH1[b * NH + c] = f2bf(d);
      }
    }
    colsum += __shfl_xor(colsum, 16, 64);
    colsum += __shfl_xor(colsum, 32, 64);
    if (g4 == 0) atomicAdd(&grad[OB1 + C * 16 + l16], colsum);
  }
  wave_sync();
  // dW1 = X^T dH1 (rows 0..11 of the 16 padded inputs)
  for (int Co = 0; Co < 4; ++Co) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < 4; ++kt) {
      const bf16x4 a = ld4s(X + (kt * 16 + 4 * g4) * KIN + l16, KIN);
      const bf16x4 b = ld4s(dH1 + (kt * 16 + 4 * g4) * NH + Co * 16 + l16, NH);
      acc = mfma16(a, b, acc);
    }
    for (int i = 0; i < 4; ++i) {
      const int f = 4 * g4 + i;
      if (f < NIN) atomicAdd(&grad[OW1 + f * NH + Co * 16 + l16], acc[i]);
    }
  }
}

// Forward only over given rows (tests, scoring).
__global__ void __launch_bounds__(64)
k_mlp_forward(const MlpDev *__restrict__ net, const float *__restrict__ Xg, int n,
              float *__restrict__ y) {
  __shared__ MlpDev s;
  __shared__ uint16_t X[64 * KIN];
  __shared__ uint16_t H1[64 * NH];
  __shared__ float ys[64];
  load_net(net, &s);
  const int lane = threadIdx.x;
  const int row = blockIdx.x * 64 + lane;
  for (int k = 0; k < KIN; ++k)
    X[lane * KIN + k] = (row < n && k < NIN) ? f2bf(Xg[(size_t)row * NIN + k]) : (uint16_t)0;
  wave_sync();
  mlp_forward_wave(&s, X, H1, nullptr, ys);
  if (row < n) y[row] = ys[lane];
}

// Adam on the FP32 master parameters, then the BF16 / FP32 kernel copies;
// clears the gradient.
__global__ void k_adam(float *__restrict__ P, float *__restrict__ G, float *__restrict__ m1,
                       float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1,
                       float c2, MlpDev *__restrict__ net) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM) return;
  const float g = G[i];
  G[i] = 0.f;
  const float a = b1 * m1[i] + (1.f - b1) * g;
  const float v = b2 * m2[i] + (1.f - b2) * g * g;
  m1[i] = a;
  m2[i] = v;
  const float w = P[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
  P[i] = w;
  if (i < OB1) net->W1[i] = f2bf(w);
  else if (i < OW2) net->b1[i - OB1] = w;
  else if (i < OB2) net->W2[i - OW2] = f2bf(w);
  else if (i < OW3) net->b2[i - OB2] = w;
  else if (i < OB3) net->w3[i - OW3] = w;
  else net->b3 = w;
}

__global__ void k_pack_net(const float *__restrict__ P, MlpDev *__restrict__ net) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < KIN * NH) net->W1[i] = (i < NIN * NH) ? f2bf(P[OW1 + i]) : (uint16_t)0;
  if (i < NH * NH) net->W2[i] = f2bf(P[OW2 + i]);
  if (i < NH) {
    net->b1[i] = P[OB1 + i];
    net->b2[i] = P[OB2 + i];
    net->w3[i] = P[OW3 + i];
  }
  if (i == 0) net->b3 = P[OB3];
}

// ---- descriptors -------------------------------------------------------------
struct NGrid {  // CSR observable grid (side >= 1.5) over one replica
  int nx, ny, nz;
  double sx, sy, sz, lx, ly, lz, hx, hy, hz;
  const int *start;
  const double *px, *py, *pz;
};

// Candidate range of one row (dx, dy) of the 27-cell stencil: the row's z
// cells hold consecutive entries of the CSR arrays, so the row is one range
// [a0, a1), plus a second [b0, b1) where the stencil wraps around z (the
// cell at the far end).  With fewer than 3 cells along z the row is every
// cell, once.
__device__ __forceinline__ void row_ranges(const NGrid &g, int base, int cz, int &a0, int &a1,
                                           int &b0, int &b1) {
  int ac0, ac1, bc0 = 0, bc1 = 0;  // cell index ranges [ac0, ac1), [bc0, bc1) within the row
  if (g.nz >= 3) {
    if (cz == 0) {
      ac0 = 0, ac1 = 2, bc0 = g.nz - 1, bc1 = g.nz;
    } else if (cz == g.nz - 1) {
      ac0 = g.nz - 2, ac1 = g.nz, bc0 = 0, bc1 = 1;
    } else {
      ac0 = cz - 1, ac1 = cz + 2;
    }
  } else {
    ac0 = 0, ac1 = g.nz;
  }
  // four loads with no branch around them (an empty second range reads one
  // entry twice), so the rows' loads all issue before any is waited for
  a0 = g.start[base + ac0];
  a1 = g.start[base + ac1];
  b0 = g.start[base + bc0];
  b1 = g.start[base + bc1];
}

// 12 smallest squared distances (ascending) from q to sphere centres in the
// 27 cells around q; entries not found stay at `far2`.  The result depends
// only on the set of candidates, not on the order they are met, so the
// search takes them row by row: the rows' ranges are loaded together, then
// each range in chunks of 8 spheres whose positions are loaded together
// (a chunk's loads are unconditional, past the range's end they repeat its
// first sphere), so a lane waits on about 20 load round trips rather than
// on one per cell and one per sphere.
__device__ __forceinline__ void take12(float d2, float best[NIN]) {
  if (d2 < best[NIN - 1]) {
    best[NIN - 1] = d2;
#pragma unroll
    for (int k = NIN - 1; k > 0; --k) {
      const float lo = fminf(best[k - 1], best[k]), hi = fmaxf(best[k - 1], best[k]);
      best[k - 1] = lo;
      best[k] = hi;
    }
  }
}

__device__ __forceinline__ void nearest12(const NGrid &g, double qx, double qy, double qz,
                                          float far2, float best[NIN]) {
  constexpr int CH = 8;
#pragma unroll
  for (int k = 0; k < NIN; ++k) best[k] = far2;
  const int cx = plain_cell(qx, g.sx, g.nx), cy = plain_cell(qy, g.sy, g.ny),
            cz = plain_cell(qz, g.sz, g.nz);
  int ra0[9], ra1[9], rb0[9], rb1[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int dx = r / 3 - 1, dy = r % 3 - 1;
    const bool ok = stencil_axis_ok(dx, g.nx) && stencil_axis_ok(dy, g.ny);
    const int base = ok ? (wrap_index(cx + dx, g.nx) * g.ny + wrap_index(cy + dy, g.ny)) * g.nz : 0;
    row_ranges(g, base, cz, ra0[r], ra1[r], rb0[r], rb1[r]);
    if (!ok) ra1[r] = ra0[r], rb1[r] = rb0[r];  // a row the stencil leaves out
  }
  auto scan = [&](int p0, int p1) {
    for (int p = p0; p < p1; p += CH) {
      double ox[CH], oy[CH], oz[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int q = p + u < p1 ? p + u : p0;
        ox[u] = g.px[q];
        oy[u] = g.py[q];
        oz[u] = g.pz[q];
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        if (p + u < p1)
          take12((float)dist2(qx, qy, qz, ox[u], oy[u], oz[u], g.lx, g.ly, g.lz, g.hx, g.hy, g.hz),
                 best);
      }
    }
  };
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    scan(ra0[r], ra1[r]);
    scan(rb0[r], rb1[r]);
  }
}

struct Vox {
  int mx, my, mz;
  double vx, vy, vz;  // voxel edges
  long V;
};

__device__ __forceinline__ void voxel_centre(const Vox &vx, long v, double &x, double &y,
                                             double &z) {
  const int iz = (int)(v % vx.mz);
  const long t = v / vx.mz;
  const int iy = (int)(t % vx.my), ix = (int)(t / vx.my);
  x = (ix + 0.5) * vx.vx;
  y = (iy + 0.5) * vx.vy;
  z = (iz + 0.5) * vx.vz;
}

// descriptor entries: distances minus one diameter
__device__ __forceinline__ void descriptor(const NGrid &g, double x, double y, double z,
                                           float far2, float out[NIN]) {
  float best[NIN];
  nearest12(g, x, y, z, far2, best);
#pragma unroll
  for (int k = 0; k < NIN; ++k) out[k] = sqrtf(best[k]) - 1.0f;
}

__global__ void __launch_bounds__(256)
k_point_descriptors(NGrid g, const double *__restrict__ q, int m, float far2,
                    float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  float d[NIN];
  descriptor(g, q[3 * i], q[3 * i + 1], q[3 * i + 2], far2, d);
  for (int k = 0; k < NIN; ++k) out[(size_t)i * NIN + k] = d[k];
}

// Training rows from uniform points of Philox stream (key, sample): input =
// descriptor of the voxel holding the point, target = -beta * max(0, 1 - d).
__global__ void __launch_bounds__(256)
k_training_rows(NGrid g, Vox vx, uint32_t k0, uint32_t k1, unsigned long long sample, long long n,
                float far2, float beta, float *__restrict__ X, float *__restrict__ T) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t u[4];
  philox4x32_10((uint32_t)i, (uint32_t)((unsigned long long)i >> 32), (uint32_t)sample,
                0x57490000u | (uint32_t)((sample >> 32) & 0xFFFFu), k0, k1, u);
  const double qx = u32_unit_closed(u[0]) * g.lx, qy = u32_unit_closed(u[1]) * g.ly,
               qz = u32_unit_closed(u[2]) * g.lz;
  float best[NIN];
  nearest12(g, qx, qy, qz, far2, best);
  const float dmin = sqrtf(best[0]);
  T[i] = -beta * fmaxf(0.f, 1.f - dmin);
  int ix = (int)(qx / vx.vx), iy = (int)(qy / vx.vy), iz = (int)(qz / vx.vz);
  ix = min(ix, vx.mx - 1);
  iy = min(iy, vx.my - 1);
  iz = min(iz, vx.mz - 1);
  const long v = ((long)ix * vx.my + iy) * vx.mz + iz;
  double cx, cy, cz;
  voxel_centre(vx, v, cx, cy, cz);
  float d[NIN];
  descriptor(g, cx, cy, cz, far2, d);
  for (int k = 0; k < NIN; ++k) X[(size_t)i * NIN + k] = d[k];
}

// Network scores of a block's 256 voxels from their descriptors (one wave per
// 64 voxels, MFMA) and the block's max and sum exp(f - max).
__device__ void score_block(const MlpDev *s, uint16_t (*X)[64 * KIN], uint16_t (*H1)[64 * NH],
                            float (*ys)[64], float *red, const float d[NIN], bool ok, long v,
                            float *__restrict__ f, float *__restrict__ bmax,
                            float *__restrict__ bsum) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = 0; k < KIN; ++k) X[w][lane * KIN + k] = k < NIN ? f2bf(d[k]) : (uint16_t)0;
  wave_sync();
  mlp_forward_wave(s, X[w], H1[w], nullptr, ys[w]);
  const float val = ys[w][lane];
  if (ok) f[v] = val;
  float m = ok ? val : -INFINITY;
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if (lane == 0) red[w] = m;
  __syncthreads();
  const float bm = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  float e = ok ? expf(val - bm) : 0.f;
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
  __syncthreads();
  if (lane == 0) red[4 + w] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    bmax[blockIdx.x] = bm;
    bsum[blockIdx.x] = red[4] + red[5] + red[6] + red[7];
  }
}

// Scores of all voxels: descriptor + network, fused (descriptors never leave
// the chip).  Block = 256 consecutive voxels, each searching the grid alone.
__global__ void __launch_bounds__(256)
k_voxel_scores(NGrid g, Vox vx, float far2, const MlpDev *__restrict__ net,
               float *__restrict__ f, float *__restrict__ bmax, float *__restrict__ bsum) {
  __shared__ MlpDev s;
  __shared__ uint16_t X[4][64 * KIN];
  __shared__ uint16_t H1[4][64 * NH];
  __shared__ float ys[4][64];
  __shared__ float red[8];
  load_net(net, &s);
  const long v = (long)blockIdx.x * 256 + threadIdx.x;
  float d[NIN];
  if (v < vx.V) {
    double cx, cy, cz;
    voxel_centre(vx, v, cx, cy, cz);
    descriptor(g, cx, cy, cz, far2, d);
  } else {
    for (int k = 0; k < NIN; ++k) d[k] = 0.f;
  }
  score_block(&s, X, H1, ys, red, d, v < vx.V, v, f, bmax, bsum);
}

// The same over 4 x 8 x 8 voxel tiles (one wave per x layer): the block first
// gathers into LDS, as FP32 offsets from the tile centre, every sphere whose
// minimum-image offset lies within the tile's half extent + smin on each
// axis (a superset of the spheres within smin of any of its voxel centres),
// then each voxel takes its 12 nearest from LDS.  The host uses it only when
// every box edge exceeds 2 (tile extent + smin), so one image serves the
// whole tile.  A tile whose set would overflow the LDS list (impossible
// below close packing: a 4 x 5 x 5 sigma^3 box holds at most ~254 spheres)
// searches the grid per voxel instead.
constexpr int TX = 4, TY = 8, TZ = 8, TCAP = 320;

__global__ void __launch_bounds__(256)
k_voxel_scores_tiled(NGrid g, Vox vx, float far2, double smin, const MlpDev *__restrict__ net,
                     float *__restrict__ f, float *__restrict__ bmax, float *__restrict__ bsum) {
  __shared__ MlpDev s;
  __shared__ uint16_t X[4][64 * KIN];
  __shared__ uint16_t H1[4][64 * NH];
  __shared__ float ys[4][64];
  __shared__ float red[8];
  __shared__ float qx[TCAP], qy[TCAP], qz[TCAP];
  __shared__ int ncand;
  if (threadIdx.x == 0) ncand = 0;
  load_net(net, &s);  // ends with a block barrier
  const int ntz = (vx.mz + TZ - 1) / TZ, nty = (vx.my + TY - 1) / TY;
  const int bz = blockIdx.x % ntz, by = (blockIdx.x / ntz) % nty, bx = blockIdx.x / (ntz * nty);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ix = bx * TX + w, iy = by * TY + (lane >> 3), iz = bz * TZ + (lane & 7);
  const bool ok = ix < vx.mx && iy < vx.my && iz < vx.mz;
  // tile centre and the half extent of the gathered box
  const double ccx = (bx * TX + TX * 0.5) * vx.vx, ccy = (by * TY + TY * 0.5) * vx.vy,
               ccz = (bz * TZ + TZ * 0.5) * vx.vz;
  const double ex = (TX - 1) * 0.5 * vx.vx + smin, ey = (TY - 1) * 0.5 * vx.vy + smin,
               ez = (TZ - 1) * 0.5 * vx.vz + smin;
  // grid cells overlapping the box, per axis (all cells when it spans them)
  int lo[3], nc[3];
  const double cc[3] = {ccx, ccy, ccz}, ee[3] = {ex, ey, ez}, side[3] = {g.sx, g.sy, g.sz};
  const int gn[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int l = (int)floor((cc[a] - ee[a]) / side[a]);
    const int h = (int)floor((cc[a] + ee[a]) / side[a]);
    nc[a] = h - l + 1;
    lo[a] = l;
    if (nc[a] >= gn[a]) {
      nc[a] = gn[a];
      lo[a] = 0;
    }
  }
  const int ncell = nc[0] * nc[1] * nc[2];
  for (int k = threadIdx.x; k < ncell; k += blockDim.x) {
    const int kx = k / (nc[1] * nc[2]), ky = (k / nc[2]) % nc[1], kz = k % nc[2];
    const int c = (wrap_index(lo[0] + kx, g.nx) * g.ny + wrap_index(lo[1] + ky, g.ny)) * g.nz +
                  wrap_index(lo[2] + kz, g.nz);
    for (int p = g.start[c]; p < g.start[c + 1]; ++p) {
      const double dx = min_image(g.px[p] - ccx, g.lx, g.hx);
      const double dy = min_image(g.py[p] - ccy, g.ly, g.hy);
      const double dz = min_image(g.pz[p] - ccz, g.lz, g.hz);
      if (fabs(dx) <= ex && fabs(dy) <= ey && fabs(dz) <= ez) {
        const int slot = atomicAdd(&ncand, 1);
        if (slot < TCAP) {
          qx[slot] = (float)dx;
          qy[slot] = (float)dy;
          qz[slot] = (float)dz;
        }
      }
    }
  }
  __syncthreads();
  const int n = ncand;
  float d[NIN];
  if (ok) {
    const double vxc = (ix + 0.5) * vx.vx, vyc = (iy + 0.5) * vx.vy, vzc = (iz + 0.5) * vx.vz;
    if (n <= TCAP) {
      float best[NIN];
#pragma unroll
      for (int k = 0; k < NIN; ++k) best[k] = far2;
      const float ox = (float)(vxc - ccx), oy = (float)(vyc - ccy), oz = (float)(vzc - ccz);
      for (int q = 0; q < n; ++q) {
        const float ax = qx[q] - ox, ay = qy[q] - oy, az = qz[q] - oz;
        const float d2 = ax * ax + ay * ay + az * az;
        if (d2 < best[NIN - 1]) {
          best[NIN - 1] = d2;
#pragma unroll
          for (int k = NIN - 1; k > 0; --k) {
            const float l2 = fminf(best[k - 1], best[k]), h2 = fmaxf(best[k - 1], best[k]);
            best[k - 1] = l2;
            best[k] = h2;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < NIN; ++k) d[k] = sqrtf(best[k]) - 1.0f;
    } else {
      descriptor(g, vxc, vyc, vzc, far2, d);
    }
  } else {
    for (int k = 0; k < NIN; ++k) d[k] = 0.f;
  }
  const long v = ((long)ix * vx.my + iy) * vx.mz + iz;
  score_block(&s, X, H1, ys, red, d, ok, v, f, bmax, bsum);
}

// Voxels that can hold a free point.  Every point of a voxel lies within
// half its diagonal h of the voxel's centre, so a voxel whose centre is
// closer than r = 1 - h to a sphere centre lies inside that sphere's
// exclusion zone and no insertion in it can succeed.  The proposal gives such
// voxels no network mass (q_v = (1 - mix) / V, the uniform part alone, still
// > 0, so the estimate stays unbiased; their score reads -inf in
// hsmc_gpu_proposal_scores) and the network scores only the others.
//
// Two passes.  k_mark_blocked: one wave per sphere marks (a byte flag)
// every voxel whose centre's squared distance to it, in FP64 as
// compute_dist evaluates it (minimum image), is below r^2 -- so a voxel is
// marked exactly when some sphere blocks it; all marks store the same 1.
// k_free_tiles: one wave per tile of 4 x 8 x 8 voxels puts the unmarked
// ("free") voxels in the tile's own range of `slots` (256 per tile, in the
// tile's x, y, z order) with their number in cnt[tile]; k_gather_free packs
// them into one list (tile order, then that order) for k_voxel_scores_list.
// No global atomics: the tile counts are prefix-summed instead.  Scores are
// written for free voxels only; the flags tell the others apart.
struct Box3 {
  double lx, ly, lz, hx, hy, hz;
};

// First candidate voxel index for centres (i + 0.5) * v within r of p: the
// centres within r have i in the open interval (x, x + 2r/v) with
// x = (p - r)/v - 0.5, and the 8 indices from floor(x - 1e-9) cover it when
// 2r/v + 1 < 8 - 1e-6 (the host checks; 6.27 for voxels of edge 0.25), also
// where rounding puts x on the other side of an integer.  The distance test
// then decides each candidate.
__device__ __forceinline__ int voxel_first(double p, double r, double v) {
  return (int)floor((p - r) / v - 0.5 - 1e-9);
}

__device__ __forceinline__ int wrap_voxel(int i, int m) {
  return i < 0 ? i + m : (i >= m ? i - m : i);
}

// Lane t's value of v, read into every lane (t the same in all of them).
__device__ __forceinline__ double lane_value(double v, int t) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, t);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), t);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__global__ void __launch_bounds__(256)
k_mark_blocked(const double *__restrict__ x, const double *__restrict__ y,
               const double *__restrict__ z, int N, Vox vx, Box3 b, double r,
               unsigned char *__restrict__ flag) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= N) return;  // the whole wave
  const double px = x[p], py = y[p], pz = z[p];
  const int lx = voxel_first(px, r, vx.vx), ly = voxel_first(py, r, vx.vy),
            lz = voxel_first(pz, r, vx.vz);
  const double r2 = r * r;
  // 8 x 8 x 8 candidates: lane's (y, z) column, 8 x-layers.  The squared
  // distance is compute_dist's (dx^2 + dy^2) + dz^2 with the minimum image
  // per axis.  Its x part is the same in every lane, so lane k < 8 computes
  // layer k's index and dx^2 once and each layer reads them from that lane.
  const int kx = lane & 7;
  const int ixk = wrap_voxel(lx + kx, vx.mx);
  const double dxk = min_image((ixk + 0.5) * vx.vx - px, b.lx, b.hx);
  const double dx2k = dxk * dxk;
  const int iy = wrap_voxel(ly + (lane >> 3), vx.my), iz = wrap_voxel(lz + (lane & 7), vx.mz);
  const double dy = min_image((iy + 0.5) * vx.vy - py, b.ly, b.hy);
  const double dz = min_image((iz + 0.5) * vx.vz - pz, b.lz, b.hz);
  const double dy2 = dy * dy, dz2 = dz * dz;
  const long col = (long)iy * vx.mz + iz, layer = (long)vx.my * vx.mz;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const double dx2 = lane_value(dx2k, k);
    const int ix = __builtin_amdgcn_readlane(ixk, k);
    if (dx2 + dy2 + dz2 < r2) flag[(long)ix * layer + col] = 1;
  }
}

__global__ void __launch_bounds__(256)
k_free_tiles(Vox vx, const unsigned char *__restrict__ flag, int *__restrict__ slots,
             int *__restrict__ cnt, long tiles) {
  const long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= tiles) return;  // the whole wave
  const int ntz = (vx.mz + TZ - 1) / TZ, nty = (vx.my + TY - 1) / TY;
  const int bz = (int)(tile % ntz), by = (int)((tile / ntz) % nty), bx = (int)(tile / ((long)ntz * nty));
  // lane l: row (x = l / 16, y = (l / 2) % 8) of the tile, voxels z = 4 (l % 2) .. + 3,
  // so lane order then z is the tile's x, y, z order
  const int lane = threadIdx.x & 63;
  const int ix = bx * TX + (lane >> 4), iy = by * TY + ((lane >> 1) & 7), iz0 = bz * TZ + 4 * (lane & 1);
  const bool row = ix < vx.mx && iy < vx.my;
  const long v0 = ((long)ix * vx.my + iy) * vx.mz + iz0;
  unsigned char f[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) f[k] = flag[row && iz0 + k < vx.mz ? v0 + k : 0];  // no branch around the loads
  int bits = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) bits |= (row && iz0 + k < vx.mz && f[k] == 0) ? 1 << k : 0;
  const int n = __popc(bits);
  const int incl = group_inclusive_sum<64>(n);
  int *out = slots + (size_t)tile * 256 + (incl - n);
#pragma unroll
  for (int k = 0, j = 0; k < 4; ++k)
    if ((bits >> k) & 1) out[j++] = (int)(v0 + k);
  if (lane == 63) cnt[tile] = incl;
}

// Pack the tiles' free voxels into one list: tile t's cnt[t] entries go to
// list[off[t] ...] (off: exclusive prefix sum of cnt); one wave per tile.
__global__ void __launch_bounds__(256)
k_gather_free(const int *__restrict__ slots, const int *__restrict__ cnt,
              const int *__restrict__ off, int *__restrict__ list, long tiles) {
  const long t = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tiles) return;
  const int n = cnt[t], o = off[t];
  for (int k = threadIdx.x & 63; k < n; k += 64) list[o + k] = slots[(size_t)t * 256 + k];
}

// Network scores of the free voxels list[0..n) (descriptor + network, as
// k_voxel_scores does for every voxel), and per block the max and sum of
// exp(f - max) over them.
__global__ void __launch_bounds__(256)
k_voxel_scores_list(NGrid g, Vox vx, float far2, const MlpDev *__restrict__ net,
                    const int *__restrict__ list, int n, float *__restrict__ f,
                    float *__restrict__ bmax, float *__restrict__ bsum) {
  __shared__ MlpDev s;
  __shared__ uint16_t X[4][64 * KIN];
  __shared__ uint16_t H1[4][64 * NH];
  __shared__ float ys[4][64];
  __shared__ float red[8];
  load_net(net, &s);
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool ok = i < n;
  const long v = ok ? (long)list[i] : 0;
  float d[NIN];
  if (ok) {
    double cx, cy, cz;
    voxel_centre(vx, v, cx, cy, cz);
    descriptor(g, cx, cy, cz, far2, d);
  } else {
    for (int k = 0; k < NIN; ++k) d[k] = 0.f;
  }
  score_block(&s, X, H1, ys, red, d, ok, v, f, bmax, bsum);
}

// Network part of the proposal mass of a free voxel v:
// mix * softmax(f)_v = mix * exp(f_v - fmax) / sumexp, with st[0] = fmax and
// st[1] = sumexp (on the device, written by k_reduce_parts).
__device__ __forceinline__ double network_mass(const float *__restrict__ f, long v,
                                               const double *__restrict__ st, double mix) {
  return mix * exp((double)f[v] - st[0]) / st[1];
}

// st[0] = max of the blocks' maxima, st[1] = sum of their sums of
// exp(f - max) rescaled to st[0], each in a fixed order (strided partials,
// then a tree); st[2] = 0.
__global__ void __launch_bounds__(256)
k_reduce_parts(const float *__restrict__ bmax, const float *__restrict__ bsum, long nb,
               double *__restrict__ st) {
  __shared__ double s[256];
  double m = -INFINITY;
  for (long b = threadIdx.x; b < nb; b += 256) m = fmax(m, (double)bmax[b]);
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  const double fm = s[0];
  __syncthreads();
  double e = 0.0;
  for (long b = threadIdx.x; b < nb; b += 256) e += (double)bsum[b] * exp((double)bmax[b] - fm);
  s[threadIdx.x] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    st[0] = fm;
    st[1] = s[0];
    st[2] = 0.0;
  }
}

// The statistics of a proposal with no free voxel (every voxel keeps the
// uniform part alone).
__global__ void k_empty_stats(double *__restrict__ st) {
  if (threadIdx.x == 0) {
    st[0] = 0.0;
    st[1] = 1.0;
    st[2] = 0.0;
  }
}

// Cumulative network mass over the free-voxel list (list order), in three
// passes: inclusive sums within blocks of 4096 entries (+ block totals), an
// exclusive scan of the block totals (one workgroup; its last entry is the
// total), then the block offsets added.
__global__ void __launch_bounds__(1024)
k_cdf_local(const float *__restrict__ f, const int *__restrict__ list, long V,
            const double *__restrict__ st, double mix, double *__restrict__ cdf,
            double *__restrict__ bsum) {
  __shared__ double s[1024];
  const long base = (long)blockIdx.x * 4096;
  double v[4], tot = 0.0;
  for (int j = 0; j < 4; ++j) {
    const long i = base + threadIdx.x * 4 + j;
    v[j] = i < V ? network_mass(f, list[i], st, mix) : 0.0;
    tot += v[j];
  }
  s[threadIdx.x] = tot;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const double t = (int)threadIdx.x >= d ? s[threadIdx.x - d] : 0.0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  double run = s[threadIdx.x] - tot;
  for (int j = 0; j < 4; ++j) {
    const long i = base + threadIdx.x * 4 + j;
    run += v[j];
    if (i < V) cdf[i] = run;
  }
  if (threadIdx.x == 1023) bsum[blockIdx.x] = s[1023];
}

__global__ void __launch_bounds__(1024) k_cdf_blocks(double *__restrict__ b, int nb) {
  __shared__ double part[1024];
  const int t = threadIdx.x, per = (nb + 1023) / 1024, lo = t * per, hi = min(nb, lo + per);
  double sum = 0.0;
  for (int i = lo; i < hi; ++i) sum += b[i];
  part[t] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const double v = t >= d ? part[t - d] : 0.0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  double run = part[t] - sum;
  for (int i = lo; i < hi; ++i) {
    const double bi = b[i];
    b[i] = run;
    run += bi;
  }
  __syncthreads();
  if (t == 1023) b[nb] = part[1023];
}

__global__ void k_cdf_add(double *__restrict__ cdf, long V, const double *__restrict__ boff) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < V) cdf[i] += boff[i / 4096];
}

// No sphere centre closer than 1 to q in the 27 cells around it (the
// overlap grid's cells are at least 1 wide).  The answer does not depend
// on the order of the cells, so they are visited nearest first (own cell,
// then the 6 sharing a face, with their ranges loaded together, then the
// rest): almost every proposed point overlaps a sphere, and usually one of
// those.
__device__ __forceinline__ bool widom_free(const NGrid &g1, double qx, double qy, double qz) {
  const int cx = plain_cell(qx, g1.sx, g1.nx), cy = plain_cell(qy, g1.sy, g1.ny),
            cz = plain_cell(qz, g1.sz, g1.nz);
  auto cell_of = [&](int j, bool &ok) {
    int dx, dy, dz;
    stencil_by_distance(j, dx, dy, dz);
    ok = stencil_axis_ok(dx, g1.nx) && stencil_axis_ok(dy, g1.ny) && stencil_axis_ok(dz, g1.nz);
    return ok ? (wrap_index(cx + dx, g1.nx) * g1.ny + wrap_index(cy + dy, g1.ny)) * g1.nz +
                    wrap_index(cz + dz, g1.nz)
              : 0;
  };
  auto hit = [&](int b, int e) {
    for (int p = b; p < e; ++p)
      if (dist2(qx, qy, qz, g1.px[p], g1.py[p], g1.pz[p], g1.lx, g1.ly, g1.lz, g1.hx, g1.hy,
                g1.hz) < 1.0)
        return true;
    return false;
  };
  int c7[7], b7[7], e7[7];
  bool ok7[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) c7[j] = cell_of(j, ok7[j]);
#pragma unroll
  for (int j = 0; j < 7; ++j) b7[j] = g1.start[c7[j]];
#pragma unroll
  for (int j = 0; j < 7; ++j) e7[j] = g1.start[c7[j] + 1];
#pragma unroll
  for (int j = 0; j < 7; ++j)
    if (!ok7[j]) e7[j] = b7[j];
#pragma unroll
  for (int j = 0; j < 7; ++j)
    if (hit(b7[j], e7[j])) return false;
  for (int j = 7; j < 27; ++j) {
    bool ok;
    const int c = cell_of(j, ok);
    if (ok && hit(g1.start[c], g1.start[c + 1])) return false;
  }
  return true;
}

// Systematic sampling of n insertions from the proposal q_v = (1 - mix) / V
// + mix softmax(f)_v (the network part on free voxels only).  Insertion i
// takes t = (i + u0) / n * T, T = (1 - mix) + *total (the network mass
// summed over the free list, mix up to rounding; 0 with no free voxel).  t below 1 - mix
// falls in the uniform part, voxel floor(t V / (1 - mix)); above, in the
// network part: the free voxel whose cumulative-mass interval holds
// t - (1 - mix).  Each voxel thus covers (1 - mix) / V + its network mass
// of the t axis and gets floor or ceil of n q_v insertions, n q_v on
// average.  A point is uniform inside its voxel and, on a hit, weighs
// 1 / (V n q_v) with q_v its whole mass over T (the flags tell free voxels
// from blocked ones, whose network mass is 0).
__global__ void __launch_bounds__(256)
k_insert(NGrid g1, Vox vx, const double *__restrict__ cdf, const int *__restrict__ list,
         long nlist, const double *__restrict__ total, long n, double u0,
         const float *__restrict__ f, const unsigned char *__restrict__ flag,
         const double *__restrict__ st, double mix, uint32_t k0, uint32_t k1,
         unsigned long long sample, double *__restrict__ part,
         unsigned long long *__restrict__ hits) {
  __shared__ double rs[4];
  __shared__ unsigned long long rh[4];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const double un = 1.0 - mix, T = un + *total;
  double w = 0.0;
  unsigned long long h = 0;
  if (i < n) {
    const double t = ((double)i + u0) / (double)n * T;
    long v;
    if (t < un || nlist == 0) {
      v = (long)(t / un * (double)vx.V);
      if (v > vx.V - 1) v = vx.V - 1;
    } else {
      const double tn = t - un;
      long lo = 0, hi = nlist - 1;  // first free voxel whose cumulative mass exceeds tn
      while (lo < hi) {
        const long mid = (lo + hi) >> 1;
        if (cdf[mid] > tn)
          hi = mid;
        else
          lo = mid + 1;
      }
      v = list[lo];
    }
    const int iz = (int)(v % vx.mz);
    const long tv = v / vx.mz;
    const int iy = (int)(tv % vx.my), ix = (int)(tv / vx.my);
  uint32_t u[4];
    philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), (uint32_t)sample,
                  0x4C490000u | (uint32_t)((sample >> 32) & 0xFFFFu), k0, k1, u);
    const double qx = (ix + ((double)u[0] + 0.5) * 0x1p-32) * vx.vx;
    const double qy = (iy + ((double)u[1] + 0.5) * 0x1p-32) * vx.vy;
    const double qz = (iz + ((double)u[2] + 0.5) * 0x1p-32) * vx.vz;
    if (widom_free(g1, qx, qy, qz)) {
      const double mass = un / (double)vx.V + (flag[v] == 0 ? network_mass(f, v, st, mix) : 0.0);
      w = 1.0 / ((double)vx.V * (double)n * (mass / T));
      h = 1;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    w += __shfl_xor(w, o, 64);
    h += __shfl_xor(h, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    rs[threadIdx.x >> 6] = w;
    rh[threadIdx.x >> 6] = h;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // the block's weight goes to its own slot and is summed in a fixed order
    // afterwards (k_sum_fixed), so an estimate never depends on block timing
    part[blockIdx.x] = rs[0] + rs[1] + rs[2] + rs[3];
    const unsigned long long sh = rh[0] + rh[1] + rh[2] + rh[3];
    if (sh) atomicAdd(hits, sh);
  }
}

// out = sum of part[0..n) in an order fixed by n alone: 256 strided partial
// sums, then a pairwise tree.
__global__ void __launch_bounds__(256) k_sum_fixed(const double *__restrict__ part, long n,
                                                   double *__restrict__ out) {
  __shared__ double s[256];
  double a = 0.0;
  for (long i = threadIdx.x; i < n; i += 256) a += part[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// Build a CSR grid of one replica (side >= cut) into the given scratch slots.
int build_ngrid(hsmc_gpu_ctx *ctx, int replica, double cut, int slot0, NGrid &g);
int build_ngrid_pair(hsmc_gpu_ctx *ctx, int replica, double cut_a, double cut_b, NGrid &ga,
                     NGrid &gb);

}  // namespace

// ---- network object ----------------------------------------------------------
struct hsmc_gpu_mlp {
  hsmc_gpu_ctx *ctx = nullptr;
  float *P = nullptr, *G = nullptr, *m1 = nullptr, *m2 = nullptr;  // FP32 master, grad, Adam
  MlpDev *net = nullptr;
  float *loss = nullptr;
  long step = 0;
  // the proposal of the last configuration scored (replica, version, step,
  // voxel); its scores, flags, free list and softmax statistics stay in the
  // context's scratch
  int sc_replica = -1;
  unsigned long long sc_version = 0;
  long sc_step = -1;
  double sc_voxel = 0.0;
  double sc_mix = -1.0;   // mix of the cached cumulative network mass
  long sc_nfree = 0;      // free voxels of the cached proposal
  // the context's score stamp when this network last wrote the shared score
  // and cumulative-mass buffers; another writer (a second network, or
  // voxel_scores) changes the context's stamp and so invalidates the cache
  unsigned long long sc_stamp = 0;
};

namespace {

__global__ void k_grid_count2(const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, int n, int nx, int ny, int nz,
                              double sx, double sy, double sz, int *__restrict__ cnt,
                              int *__restrict__ cellof) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (plain_cell(x[i], sx, nx) * ny + plain_cell(y[i], sy, ny)) * nz + plain_cell(z[i], sz, nz);
  cellof[i] = c;
  atomicAdd(&cnt[c], 1);
}

__global__ void k_grid_scatter2(const double *__restrict__ x, const double *__restrict__ y,
                                const double *__restrict__ z, int n, const int *__restrict__ cellof,
                                const int *__restrict__ start, int *__restrict__ fill,
                                double *__restrict__ px, double *__restrict__ py,
                                double *__restrict__ pz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cellof[i];
  const int pos = start[c] + atomicAdd(&fill[c], 1);
  px[pos] = x[i];
  py[pos] = y[i];
  pz[pos] = z[i];
}

// Two CSR grids of one replica filed in one pass: grid a's cells are the
// entries [0, Ca) of cnt and grid b's the entries [Ca, Ca + Cb), so one
// prefix sum over both gives grid a's starts in [0, n] and grid b's in
// [n, 2n], and both grids index one position array of 2n entries.
struct CellDims {
  int nx, ny, nz;
  double sx, sy, sz;
};

__device__ __forceinline__ int cell_in(const CellDims &d, double x, double y, double z) {
  return (plain_cell(x, d.sx, d.nx) * d.ny + plain_cell(y, d.sy, d.ny)) * d.nz +
         plain_cell(z, d.sz, d.nz);
}

__global__ void k_grid_count_pair(const double *__restrict__ x, const double *__restrict__ y,
                                  const double *__restrict__ z, int n, CellDims a, CellDims b,
                                  int ca, int *__restrict__ cnt, int *__restrict__ cellof) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x[i], yi = y[i], zi = z[i];
  const int c0 = cell_in(a, xi, yi, zi), c1 = ca + cell_in(b, xi, yi, zi);
  cellof[i] = c0;
  cellof[n + i] = c1;
  atomicAdd(&cnt[c0], 1);
  atomicAdd(&cnt[c1], 1);
}

__global__ void k_grid_scatter_pair(const double *__restrict__ x, const double *__restrict__ y,
                                    const double *__restrict__ z, int n,
                                    const int *__restrict__ cellof, const int *__restrict__ start,
                                    int *__restrict__ fill, double *__restrict__ px,
                                    double *__restrict__ py, double *__restrict__ pz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x[i], yi = y[i], zi = z[i];
  const int c0 = cellof[i], c1 = cellof[n + i];
  const int s0 = start[c0], s1 = start[c1];
  const int p0 = s0 + atomicAdd(&fill[c0], 1), p1 = s1 + atomicAdd(&fill[c1], 1);
  px[p0] = xi;
  py[p0] = yi;
  pz[p0] = zi;
  px[p1] = xi;
  py[p1] = yi;
  pz[p1] = zi;
}

int grid_cells_for2(double L, double cut) {
  int n = (int)floor(L / cut);
  if (n < 1) n = 1;
  while (n > 1 && L / n < cut) --n;
  return n;
}

// scratch slots 24..43 belong to this file (observables.hip uses 0..13,
// ctx.hip 21..22, sweep.hip 23): descriptor grid 24..26, the insertion call's
// descriptor and overlap grids built together 27..29,
// rows / scores 30, targets / block partials 31, cumulative mass 32,
// sums 34, per-block insertion weights 35, cumulative-mass block sums 36,
// grid-scan partials 37, voxel scores 38, score block reductions 39,
// free-voxel tile slots and list 40, tile counts, offsets and scan space 41,
// blocked-voxel flags 42, softmax statistics 43.
enum { L_GRID_A = 24, L_GRID_PAIR = 27, L_X = 30, L_T = 31, L_CDF = 32, L_ACC = 34, L_PART = 35,
       L_BSUM = 36, L_SCANP = 37, L_F = 38, L_FRED = 39, L_LIST = 40, L_NFREE = 41, L_FLAG = 42,
       L_STATS = 43 };

void set_grid(NGrid &g, const RepHost &h, double cut, const int *start, const double *pos,
              size_t stride) {
  g.nx = grid_cells_for2(h.p.lx, cut);
  g.ny = grid_cells_for2(h.p.ly, cut);
  g.nz = grid_cells_for2(h.p.lz, cut);
  g.sx = h.p.lx / g.nx; g.sy = h.p.ly / g.ny; g.sz = h.p.lz / g.nz;
  g.lx = h.p.lx; g.ly = h.p.ly; g.lz = h.p.lz;
  g.hx = h.p.lx / 2.0; g.hy = h.p.ly / 2.0; g.hz = h.p.lz / 2.0;
  g.start = start;
  g.px = pos; g.py = pos + stride; g.pz = pos + 2 * stride;
}

// Descriptor grid (cut_a) and overlap grid (cut_b) of `replica` in one set
// of launches: the cache entry 1 of the context's grid cache, keyed by
// replica and configuration version (the cuts are the insertion call's
// constants).
int build_ngrid_pair(hsmc_gpu_ctx *ctx, int replica, double cut_a, double cut_b, NGrid &ga,
                     NGrid &gb) {
  const RepHost &h = ctx->reps[replica];
  const int N = h.p.n_part;
  const long Ca = (long)grid_cells_for2(h.p.lx, cut_a) * grid_cells_for2(h.p.ly, cut_a) *
                  grid_cells_for2(h.p.lz, cut_a);
  const long Cb = (long)grid_cells_for2(h.p.lx, cut_b) * grid_cells_for2(h.p.ly, cut_b) *
                  grid_cells_for2(h.p.lz, cut_b);
  const long C = Ca + Cb;
  if (C + 1 > (1L << 31) - 1 || 2L * N > (1L << 31) - 1) return HSMC_ERR_NOT_SUPPORTED;
  const bool cached = ctx->ngrid_replica[1] == replica && ctx->ngrid_version[1] == ctx->version &&
                      ctx->ngrid_cut[1] == cut_a + cut_b;
  int *cnt = (int *)scratch(ctx, L_GRID_PAIR + 0, sizeof(int) * 2 * (C + 1) + sizeof(int) * 2 * N);
  double *pos = (double *)scratch(ctx, L_GRID_PAIR + 1, sizeof(double) * 6 * (size_t)N);
  int *start = (int *)scratch(ctx, L_GRID_PAIR + 2, sizeof(int) * (C + 1));
  int *part = (int *)scratch(ctx, L_SCANP, sizeof(int) * scan_workspace_ints(C));
  if (!cnt || !pos || !start || !part) return HSMC_ERR_NO_MEMORY;
  int *fill = cnt + (C + 1);
  int *cellof = fill + (C + 1);
  const size_t stride = 2 * (size_t)N;
  set_grid(ga, h, cut_a, start, pos, stride);
  set_grid(gb, h, cut_b, start + Ca, pos, stride);
  if (cached) return HSMC_OK;
  const CellDims da{ga.nx, ga.ny, ga.nz, ga.sx, ga.sy, ga.sz};
  const CellDims db{gb.nx, gb.ny, gb.nz, gb.sx, gb.sy, gb.sz};
  const double *x = ctx->d_x + h.poff, *y = ctx->d_y + h.poff, *z = ctx->d_z + h.poff;
  HSMC_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int) * 2 * (C + 1), ctx->stream));
  hipLaunchKernelGGL(k_grid_count_pair, dim3((N + 255) / 256), dim3(256), 0, ctx->stream, x, y, z,
                     N, da, db, (int)Ca, cnt, cellof);
  exclusive_scan(ctx->stream, cnt, (int)C, start, part);
  hipLaunchKernelGGL(k_grid_scatter_pair, dim3((N + 255) / 256), dim3(256), 0, ctx->stream, x, y,
                     z, N, cellof, start, fill, pos, pos + stride, pos + 2 * stride);
  HSMC_HIP(ctx, hipGetLastError());
  ctx->ngrid_replica[1] = replica;
  ctx->ngrid_version[1] = ctx->version;
  ctx->ngrid_cut[1] = cut_a + cut_b;
  return HSMC_OK;
}

int build_ngrid(hsmc_gpu_ctx *ctx, int replica, double cut, int slot0, NGrid &g) {
  const RepHost &h = ctx->reps[replica];
  const int N = h.p.n_part;
  const int nx = grid_cells_for2(h.p.lx, cut), ny = grid_cells_for2(h.p.ly, cut),
            nz = grid_cells_for2(h.p.lz, cut);
  const long C = (long)nx * ny * nz;
  const int which = 0;  // entry 1 is the paired build's
  const bool cached = ctx->ngrid_replica[which] == replica &&
                      ctx->ngrid_version[which] == ctx->version && ctx->ngrid_cut[which] == cut;
  int *cnt = (int *)scratch(ctx, slot0 + 0, sizeof(int) * 2 * (C + 1) + sizeof(int) * N);
  double *pos = (double *)scratch(ctx, slot0 + 1, sizeof(double) * 3 * N);
  int *start = (int *)scratch(ctx, slot0 + 2, sizeof(int) * (C + 1));
  int *part = (int *)scratch(ctx, L_SCANP, sizeof(int) * scan_workspace_ints(C));
  if (!cnt || !pos || !start || !part) return HSMC_ERR_NO_MEMORY;
  int *fill = cnt + (C + 1);
  int *cellof = fill + (C + 1);
  const double *x = ctx->d_x + h.poff, *y = ctx->d_y + h.poff, *z = ctx->d_z + h.poff;
  const double sx = h.p.lx / nx, sy = h.p.ly / ny, sz = h.p.lz / nz;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.sx = sx; g.sy = sy; g.sz = sz;
  g.lx = h.p.lx; g.ly = h.p.ly; g.lz = h.p.lz;
  g.hx = h.p.lx / 2.0; g.hy = h.p.ly / 2.0; g.hz = h.p.lz / 2.0;
  g.start = start;
  g.px = pos; g.py = pos + N; g.pz = pos + 2 * (size_t)N;
  if (cached) return HSMC_OK;
  HSMC_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int) * 2 * (C + 1), ctx->stream));
  hipLaunchKernelGGL(k_grid_count2, dim3((N + 255) / 256), dim3(256), 0, ctx->stream, x, y, z, N,
                     nx, ny, nz, sx, sy, sz, cnt, cellof);
  exclusive_scan(ctx->stream, cnt, (int)C, start, part);
  hipLaunchKernelGGL(k_grid_scatter2, dim3((N + 255) / 256), dim3(256), 0, ctx->stream, x, y, z,
                     N, cellof, start, fill, pos, pos + N, pos + 2 * (size_t)N);
  HSMC_HIP(ctx, hipGetLastError());
  ctx->ngrid_replica[which] = replica;
  ctx->ngrid_version[which] = ctx->version;
  ctx->ngrid_cut[which] = cut;
  return HSMC_OK;
}

constexpr double kDescCut = 1.5;  // descriptor grid side (sigma)

Vox make_vox(const RepHost &h, double voxel) {
  Vox v;
  v.mx = (int)ceil(h.p.lx / voxel);
  v.my = (int)ceil(h.p.ly / voxel);
  v.mz = (int)ceil(h.p.lz / voxel);
  v.vx = h.p.lx / v.mx;
  v.vy = h.p.ly / v.my;
  v.vz = h.p.lz / v.mz;
  v.V = (long)v.mx * v.my * v.mz;
  return v;
}

float far2_of(const NGrid &g) {
  const double s = std::min(g.sx, std::min(g.sy, g.sz));
  return (float)(s * s);
}

// Voxel scores into f, and per-block max / sum exp(f - max) into the scratch
// returned in *bred (*nblk blocks): the tiled kernel when every box edge
// exceeds 2 (tile extent + smin), else the per-voxel one.
int launch_scores(hsmc_gpu_ctx *ctx, const NGrid &g, const Vox &vx, const MlpDev *net, float *f,
                  float **bred, long *nblk) {
  const double smin = std::min(g.sx, std::min(g.sy, g.sz));
  const bool tiled = g.lx > 2.0 * ((TX - 1) * vx.vx + smin) * (1.0 + 1e-9) &&
                     g.ly > 2.0 * ((TY - 1) * vx.vy + smin) * (1.0 + 1e-9) &&
                     g.lz > 2.0 * ((TZ - 1) * vx.vz + smin) * (1.0 + 1e-9);
  const long nb = tiled ? (long)((vx.mx + TX - 1) / TX) * ((vx.my + TY - 1) / TY) * ((vx.mz + TZ - 1) / TZ)
                        : (vx.V + 255) / 256;
  float *b = (float *)scratch(ctx, L_FRED, sizeof(float) * 2 * (size_t)nb);
  if (!b) return HSMC_ERR_NO_MEMORY;
  if (tiled)
    hipLaunchKernelGGL(k_voxel_scores_tiled, dim3((unsigned)nb), dim3(256), 0, ctx->stream, g, vx,
                       far2_of(g), smin, net, f, b, b + nb);
  else
    hipLaunchKernelGGL(k_voxel_scores, dim3((unsigned)nb), dim3(256), 0, ctx->stream, g, vx,
                       far2_of(g), net, f, b, b + nb);
  HSMC_HIP(ctx, hipGetLastError());
  *bred = b;
  *nblk = nb;
  return HSMC_OK;
}

__global__ void k_iota(int *__restrict__ list, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) list[i] = (int)i;
}

// -inf for the voxels the flags mark as blocked (the proposal's scores as a
// caller sees them).
__global__ void k_mask_blocked(float *__restrict__ f, const unsigned char *__restrict__ flag,
                               long V) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < V && flag[i]) f[i] = -INFINITY;
}

// The proposal of one configuration: the byte flags of blocked voxels
// (k_mark_blocked), the list of free voxels (k_free_tiles, k_gather_free),
// the network's scores of the free voxels in f (k_voxel_scores_list), and
// in stats[0], stats[1] (device) fmax and sumexp over them (0 and 1 when none is
// free, so every voxel keeps the uniform part alone).  Where a voxel is too
// large for any exclusion, or too small for the marking kernel's 8
// candidates per axis, every voxel is free and scored.  Only the number of
// free voxels comes back to the host (it sizes the later launches).
struct Proposal {
  unsigned char *flag = nullptr;
  int *list = nullptr;
  double *stats = nullptr;
  long nfree = 0;
};

int launch_proposal_scores(hsmc_gpu_ctx *ctx, int replica, const NGrid &g, const Vox &vx,
                           const MlpDev *net, float *f, Proposal *pr) {
  const double h = 0.5 * sqrt(vx.vx * vx.vx + vx.vy * vx.vy + vx.vz * vx.vz);
  const double r = 1.0 - h - 1e-9;
  // the marking kernel's 8 candidate voxels per axis must cover the centres
  // within r of a sphere (voxel_first), and each axis must have 8 voxels to
  // wrap over
  const double lim = 8.0 - 1e-6;
  const bool span_ok = 2.0 * r / vx.vx + 1.0 < lim && 2.0 * r / vx.vy + 1.0 < lim &&
                       2.0 * r / vx.vz + 1.0 < lim && vx.mx >= 8 && vx.my >= 8 && vx.mz >= 8;
  const long tiles = (long)((vx.mx + TX - 1) / TX) * ((vx.my + TY - 1) / TY) * ((vx.mz + TZ - 1) / TZ);
  // per tile: 256 free-voxel slots, then the packed list; the tile counts,
  // their prefix sum and the scan's workspace; a byte per voxel
  int *slots = (int *)scratch(ctx, L_LIST, sizeof(int) * ((size_t)tiles * 256 + (size_t)vx.V));
  int *cnt = (int *)scratch(ctx, L_NFREE,
                            sizeof(int) * (2 * (size_t)tiles + 1 + (size_t)scan_workspace_ints(tiles)));
  unsigned char *flag = (unsigned char *)scratch(ctx, L_FLAG, (size_t)vx.V);
  double *stats = (double *)scratch(ctx, L_STATS, sizeof(double) * 3);
  if (!slots || !cnt || !flag || !stats) return HSMC_ERR_NO_MEMORY;
  int *list = slots + (size_t)tiles * 256;
  pr->flag = flag;
  pr->list = list;
  pr->stats = stats;
  HSMC_HIP(ctx, hipMemsetAsync(flag, 0, (size_t)vx.V, ctx->stream));
  if (!(r > 0.0) || !span_ok) {
    float *bred = nullptr;
    long nblk = 0;
    int st = launch_scores(ctx, g, vx, net, f, &bred, &nblk);
    if (st) return st;
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((vx.V + 255) / 256)), dim3(256), 0, ctx->stream, list,
                       vx.V);
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, ctx->stream, bred, bred + nblk, nblk,
                       stats);
    HSMC_HIP(ctx, hipGetLastError());
    pr->nfree = vx.V;
    return HSMC_OK;
  }
  const RepHost &rh = ctx->reps[replica];
  int *off = cnt + tiles;
  int *part = off + tiles + 1;
  const int N = rh.p.n_part;
  Box3 b3{rh.p.lx, rh.p.ly, rh.p.lz, rh.p.lx / 2.0, rh.p.ly / 2.0, rh.p.lz / 2.0};
  hipLaunchKernelGGL(k_mark_blocked, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, ctx->stream,
                     ctx->d_x + rh.poff, ctx->d_y + rh.poff, ctx->d_z + rh.poff, N, vx, b3, r, flag);
  hipLaunchKernelGGL(k_free_tiles, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, ctx->stream, vx,
                     flag, slots, cnt, tiles);
  exclusive_scan(ctx->stream, cnt, (int)tiles, off, part);
  hipLaunchKernelGGL(k_gather_free, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, ctx->stream,
                     slots, cnt, off, list, tiles);
  HSMC_HIP(ctx, hipGetLastError());
  int nf = 0;
  HSMC_HIP(ctx, hipMemcpyAsync(&nf, off + tiles, sizeof(nf), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  pr->nfree = nf;
  if (nf == 0) {
    hipLaunchKernelGGL(k_empty_stats, dim3(1), dim3(64), 0, ctx->stream, stats);
    HSMC_HIP(ctx, hipGetLastError());
    return HSMC_OK;
  }
  const long nb = ((long)nf + 255) / 256;
  float *b = (float *)scratch(ctx, L_FRED, sizeof(float) * 2 * (size_t)nb);
  if (!b) return HSMC_ERR_NO_MEMORY;
  // the softmax totals from the scoring kernel's per-block max and sum of
  // exp(f - max) (each block's 256 list entries reduced in a fixed order),
  // combined in a fixed order: they depend only on the scores and the list
  hipLaunchKernelGGL(k_voxel_scores_list, dim3((unsigned)nb), dim3(256), 0, ctx->stream, g, vx,
                     far2_of(g), net, list, (int)nf, f, b, b + nb);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, ctx->stream, b, b + nb, nb, stats);
  HSMC_HIP(ctx, hipGetLastError());
  return HSMC_OK;
}

}  // namespace

extern "C" {

int hsmc_gpu_mlp_create(hsmc_gpu_ctx *ctx, hsmc_gpu_mlp **out, uint64_t seed) {
  if (!ctx || !out) return HSMC_ERR_BAD_ARG;
  hsmc_gpu_mlp *m = new hsmc_gpu_mlp();
  m->ctx = ctx;
  if (hipMalloc(&m->P, sizeof(float) * NPARAM) != hipSuccess ||
      hipMalloc(&m->G, sizeof(float) * NPARAM) != hipSuccess ||
      hipMalloc(&m->m1, sizeof(float) * NPARAM) != hipSuccess ||
      hipMalloc(&m->m2, sizeof(float) * NPARAM) != hipSuccess ||
      hipMalloc(&m->net, sizeof(MlpDev)) != hipSuccess ||
      hipMalloc(&m->loss, sizeof(float)) != hipSuccess) {
    hsmc_gpu_mlp_destroy(m);
    return HSMC_ERR_NO_MEMORY;
  }
  // He-uniform initial weights from Philox(seed), zero biases
  std::vector<float> p(NPARAM, 0.f);
  auto uni = [&](uint32_t i) {
    uint32_t u[4];
    philox4x32_10(i, 0x4D4C5000u, 0u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), u);
    return (float)(((double)u[0] + 0.5) * 0x1p-32 * 2.0 - 1.0);
  };
  const float a1 = sqrtf(6.f / NIN), a2 = sqrtf(6.f / NH);
  for (int i = 0; i < NIN * NH; ++i) p[OW1 + i] = a1 * uni(i);
  for (int i = 0; i < NH * NH; ++i) p[OW2 + i] = a2 * uni(10000 + i);
  for (int i = 0; i < NH; ++i) p[OW3 + i] = a2 * uni(20000 + i);
  int st = hsmc_gpu_mlp_set(m, p.data());
  if (st) {
    hsmc_gpu_mlp_destroy(m);
    return st;
  }
  *out = m;
  return HSMC_OK;
}

int hsmc_gpu_mlp_destroy(hsmc_gpu_mlp *m) {
  if (!m) return HSMC_OK;
  if (m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
  void *b[] = {m->P, m->G, m->m1, m->m2, m->net, m->loss};
  for (void *x : b)
    if (x) (void)hipFree(x);
  delete m;
  return HSMC_OK;
}

int hsmc_gpu_mlp_forget(hsmc_gpu_mlp *m) {
  if (!m) return HSMC_ERR_BAD_ARG;
  m->sc_replica = -1;
  return HSMC_OK;
}

int hsmc_gpu_mlp_set(hsmc_gpu_mlp *m, const float *params) {
  if (!m || !params) return HSMC_ERR_BAD_ARG;
  hsmc_gpu_ctx *ctx = m->ctx;
  HSMC_HIP(ctx, hipMemcpyAsync(m->P, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(m->G, 0, sizeof(float) * NPARAM, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(m->m1, 0, sizeof(float) * NPARAM, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(m->m2, 0, sizeof(float) * NPARAM, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(m->net, 0, sizeof(MlpDev), ctx->stream));
  hipLaunchKernelGGL(k_pack_net, dim3((NH * NH + 255) / 256), dim3(256), 0, ctx->stream, m->P, m->net);
  HSMC_HIP(ctx, hipGetLastError());
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  m->step = 0;
  m->sc_replica = -1;
  return HSMC_OK;
}

int hsmc_gpu_mlp_get(hsmc_gpu_mlp *m, float *params) {
  if (!m || !params) return HSMC_ERR_BAD_ARG;
  hsmc_gpu_ctx *ctx = m->ctx;
  HSMC_HIP(ctx, hipMemcpyAsync(params, m->P, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HSMC_OK;
}

int hsmc_gpu_mlp_forward(hsmc_gpu_mlp *m, const float *x, int n, float *y) {
  if (!m || n < 0 || (n > 0 && (!x || !y))) return HSMC_ERR_BAD_ARG;
  if (n == 0) return HSMC_OK;
  hsmc_gpu_ctx *ctx = m->ctx;
  float *dx = (float *)scratch(ctx, L_X, sizeof(float) * NIN * (size_t)n);
  float *dy = (float *)scratch(ctx, L_T, sizeof(float) * (size_t)n);
  if (!dx || !dy) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemcpyAsync(dx, x, sizeof(float) * NIN * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_mlp_forward, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, m->net, dx, n, dy);
  HSMC_HIP(ctx, hipGetLastError());
  HSMC_HIP(ctx, hipMemcpyAsync(y, dy, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HSMC_OK;
}

int hsmc_gpu_mlp_gradient(hsmc_gpu_mlp *m, const float *x, const float *t, int n, float *loss,
                          float *grad) {
  if (!m || n <= 0 || !x || !t) return HSMC_ERR_BAD_ARG;
  hsmc_gpu_ctx *ctx = m->ctx;
  float *dx = (float *)scratch(ctx, L_X, sizeof(float) * NIN * (size_t)n);
  float *dt = (float *)scratch(ctx, L_T, sizeof(float) * (size_t)n);
  if (!dx || !dt) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemcpyAsync(dx, x, sizeof(float) * NIN * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(dt, t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(m->G, 0, sizeof(float) * NPARAM, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(m->loss, 0, sizeof(float), ctx->stream));
  hipLaunchKernelGGL(k_mlp_train_step, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, m->net, dx,
                     dt, n, 1.0f / n, m->G, m->loss, nullptr);
  HSMC_HIP(ctx, hipGetLastError());
  float l = 0.f;
  HSMC_HIP(ctx, hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  if (grad)
    HSMC_HIP(ctx, hipMemcpyAsync(grad, m->G, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(m->G, 0, sizeof(float) * NPARAM, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (loss) *loss = l / n;
  return HSMC_OK;
}

int hsmc_gpu_point_descriptors(hsmc_gpu_ctx *ctx, int replica, const double *xyz, int m,
                               float *desc) {
  if (!ctx || replica < 0 || replica >= ctx->R || m < 0 || (m > 0 && (!xyz || !desc)))
    return HSMC_ERR_BAD_ARG;
  if (m == 0) return HSMC_OK;
  NGrid g;
  int st = build_ngrid(ctx, replica, kDescCut, L_GRID_A, g);
  if (st) return st;
  double *dq = (double *)scratch(ctx, L_X, sizeof(double) * 3 * (size_t)m);
  float *dd = (float *)scratch(ctx, L_T, sizeof(float) * NIN * (size_t)m);
  if (!dq || !dd) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemcpyAsync(dq, xyz, sizeof(double) * 3 * m, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_point_descriptors, dim3((m + 255) / 256), dim3(256), 0, ctx->stream, g, dq,
                     m, far2_of(g), dd);
  HSMC_HIP(ctx, hipGetLastError());
  HSMC_HIP(ctx, hipMemcpyAsync(desc, dd, sizeof(float) * NIN * m, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HSMC_OK;
}

int hsmc_gpu_mlp_train(hsmc_gpu_mlp *m, int replica, double voxel, uint64_t seed, uint64_t sample,
                       int64_t n_points, int epochs, int batch, float lr, float beta,
                       float *final_loss) {
  if (!m || n_points <= 0 || epochs <= 0 || batch < 64 || !(voxel > 0.0)) return HSMC_ERR_BAD_ARG;
  hsmc_gpu_ctx *ctx = m->ctx;
  if (replica < 0 || replica >= ctx->R) return HSMC_ERR_BAD_ARG;
  NGrid g;
  int st = build_ngrid(ctx, replica, kDescCut, L_GRID_A, g);
  if (st) return st;
  const Vox vx = make_vox(ctx->reps[replica], voxel);
  float *X = (float *)scratch(ctx, L_X, sizeof(float) * NIN * (size_t)n_points);
  float *T = (float *)scratch(ctx, L_T, sizeof(float) * (size_t)n_points);
  if (!X || !T) return HSMC_ERR_NO_MEMORY;
  hipLaunchKernelGGL(k_training_rows, dim3((unsigned)((n_points + 255) / 256)), dim3(256), 0,
                     ctx->stream, g, vx, (uint32_t)seed, (uint32_t)(seed >> 32),
                     (unsigned long long)sample, (long long)n_points, far2_of(g), beta, X, T);
  HSMC_HIP(ctx, hipGetLastError());
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float last = 0.f;
  for (int e = 0; e < epochs; ++e) {
    HSMC_HIP(ctx, hipMemsetAsync(m->loss, 0, sizeof(float), ctx->stream));
    for (int64_t b0 = 0; b0 < n_points; b0 += batch) {
      const int nb = (int)std::min<int64_t>(batch, n_points - b0);
      hipLaunchKernelGGL(k_mlp_train_step, dim3((nb + 63) / 64), dim3(64), 0, ctx->stream, m->net,
                         X + (size_t)b0 * NIN, T + b0, nb, 1.0f / nb, m->G, m->loss, nullptr);
      ++m->step;
      const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
      hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, m->P, m->G,
                         m->m1, m->m2, lr, b1, b2, eps, c1, c2, m->net);
    }
    HSMC_HIP(ctx, hipGetLastError());
    HSMC_HIP(ctx, hipMemcpyAsync(&last, m->loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (final_loss) *final_loss = last / (float)n_points;
  return HSMC_OK;
}

int hsmc_gpu_widom_learned(hsmc_gpu_mlp *m, int replica, double voxel, double mix, uint64_t seed,
                           uint64_t sample, int64_t n, double *estimate, int64_t *n_done,
                           int64_t *hits) {
  if (!m || !estimate || n <= 0 || n > (1LL << 30) || !(mix >= 0.0 && mix < 1.0) || !(voxel > 0.0))
    return HSMC_ERR_BAD_ARG;
  hsmc_gpu_ctx *ctx = m->ctx;
  if (replica < 0 || replica >= ctx->R) return HSMC_ERR_BAD_ARG;
  NGrid g, g1;
  int st = build_ngrid_pair(ctx, replica, kDescCut, 1.0 + 1e-9, g, g1);
  if (st) return st;
  const Vox vx = make_vox(ctx->reps[replica], voxel);
  if (vx.V > (1L << 31) - 1) return HSMC_ERR_NOT_SUPPORTED;
  float *f = (float *)scratch(ctx, L_F, sizeof(float) * (size_t)vx.V);
  double *acc = (double *)scratch(ctx, L_ACC, sizeof(double) * 2);
  const long nblk_ins = (long)((n + 255) / 256);
  double *part = (double *)scratch(ctx, L_PART, sizeof(double) * (size_t)nblk_ins);
  if (!f || !acc || !part) return HSMC_ERR_NO_MEMORY;
  const bool scores_ok = m->sc_replica == replica && m->sc_version == ctx->version &&
                         m->sc_step == m->step && m->sc_voxel == voxel &&
                         m->sc_stamp == ctx->score_stamp;
  Proposal pr;
  if (!scores_ok) {
    st = launch_proposal_scores(ctx, replica, g, vx, m->net, f, &pr);
    if (st) return st;
    m->sc_replica = replica;
    m->sc_version = ctx->version;
    m->sc_step = m->step;
    m->sc_voxel = voxel;
    m->sc_nfree = pr.nfree;
    m->sc_mix = -1.0;  // the cumulative mass is rebuilt below
    m->sc_stamp = ++ctx->score_stamp;
  }
  // the cached proposal's flags and free list (the same scratch the scoring
  // wrote; the stamp check above guarantees nobody rewrote it since)
  const long tiles = (long)((vx.mx + TX - 1) / TX) * ((vx.my + TY - 1) / TY) * ((vx.mz + TZ - 1) / TZ);
  int *slots = (int *)scratch(ctx, L_LIST, sizeof(int) * ((size_t)tiles * 256 + (size_t)vx.V));
  unsigned char *flag = (unsigned char *)scratch(ctx, L_FLAG, (size_t)vx.V);
  double *stats = (double *)scratch(ctx, L_STATS, sizeof(double) * 3);
  if (!slots || !flag || !stats) return HSMC_ERR_NO_MEMORY;
  const int *list = slots + (size_t)tiles * 256;
  const long nfree = m->sc_nfree;
  const long ncdf = std::max(1L, (nfree + 4095) / 4096);
  double *cdf = (double *)scratch(ctx, L_CDF, sizeof(double) * (size_t)std::max(1L, nfree));
  double *bsum = (double *)scratch(ctx, L_BSUM, sizeof(double) * (size_t)(ncdf + 1));
  if (!cdf || !bsum) return HSMC_ERR_NO_MEMORY;
  if (m->sc_mix != mix && nfree > 0) {
    hipLaunchKernelGGL(k_cdf_local, dim3((unsigned)ncdf), dim3(1024), 0, ctx->stream, f, list,
                       nfree, stats, mix, cdf, bsum);
    hipLaunchKernelGGL(k_cdf_blocks, dim3(1), dim3(1024), 0, ctx->stream, bsum, (int)ncdf);
    hipLaunchKernelGGL(k_cdf_add, dim3((unsigned)((nfree + 255) / 256)), dim3(256), 0, ctx->stream,
                       cdf, nfree, bsum);
    HSMC_HIP(ctx, hipGetLastError());
  }
  if (m->sc_mix != mix) {
    m->sc_mix = mix;
    m->sc_stamp = ++ctx->score_stamp;
  }
  // the network mass's total: the scan's last entry, or 0 with no free voxel
  const double *total = nfree > 0 ? bsum + ncdf : stats + 2;
  // one uniform offset per call for the systematic sample
  uint32_t u[4];
  philox4x32_10(0u, 0u, (uint32_t)sample, 0x4C530000u | (uint32_t)((sample >> 32) & 0xFFFFu),
                (uint32_t)seed, (uint32_t)(seed >> 32), u);
  const double u0 = (double)u[0] * 0x1p-32;
  HSMC_HIP(ctx, hipMemsetAsync(acc, 0, sizeof(double) * 2, ctx->stream));
  hipLaunchKernelGGL(k_insert, dim3((unsigned)nblk_ins), dim3(256), 0, ctx->stream, g1, vx,
                     cdf, list, nfree, total, (long)n, u0, f, flag, stats, mix,
                     (uint32_t)seed, (uint32_t)(seed >> 32), (unsigned long long)sample, part,
                     reinterpret_cast<unsigned long long *>(acc + 1));
  hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(256), 0, ctx->stream, part, nblk_ins, acc);
  HSMC_HIP(ctx, hipGetLastError());
  double ha[2] = {0, 0};
  HSMC_HIP(ctx, hipMemcpyAsync(ha, acc, sizeof(double) * 2, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *estimate = ha[0];
  if (n_done) *n_done = n;
  if (hits) *hits = (int64_t)__builtin_bit_cast(unsigned long long, ha[1]);
  return HSMC_OK;
}

int hsmc_gpu_voxel_scores(hsmc_gpu_mlp *m, int replica, double voxel, float *scores, int64_t cap,
                          int64_t *n_voxels) {
  if (!m || !(voxel > 0.0) || !n_voxels) return HSMC_ERR_BAD_ARG;
  hsmc_gpu_ctx *ctx = m->ctx;
  if (replica < 0 || replica >= ctx->R) return HSMC_ERR_BAD_ARG;
  const Vox vx = make_vox(ctx->reps[replica], voxel);
  *n_voxels = vx.V;
  if (!scores) return HSMC_OK;
  if (cap < vx.V) return HSMC_ERR_BAD_ARG;
  NGrid g;
  int st = build_ngrid(ctx, replica, kDescCut, L_GRID_A, g);
  if (st) return st;
  float *f = (float *)scratch(ctx, L_F, sizeof(float) * (size_t)vx.V);
  if (!f) return HSMC_ERR_NO_MEMORY;
  // the shared score buffer is overwritten: no network's cache stays valid
  ++ctx->score_stamp;
  m->sc_replica = -1;
  float *bred = nullptr;
  long nblk = 0;
  st = launch_scores(ctx, g, vx, m->net, f, &bred, &nblk);
  if (st) return st;
  HSMC_HIP(ctx, hipMemcpyAsync(scores, f, sizeof(float) * vx.V, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HSMC_OK;
}

int hsmc_gpu_proposal_scores(hsmc_gpu_mlp *m, int replica, double voxel, float *scores,
                             int64_t cap, int64_t *n_voxels, int64_t *n_free) {
  if (!m || !(voxel > 0.0) || !n_voxels) return HSMC_ERR_BAD_ARG;
  hsmc_gpu_ctx *ctx = m->ctx;
  if (replica < 0 || replica >= ctx->R) return HSMC_ERR_BAD_ARG;
  const Vox vx = make_vox(ctx->reps[replica], voxel);
  *n_voxels = vx.V;
  if (!scores) return HSMC_OK;
  if (cap < vx.V) return HSMC_ERR_BAD_ARG;
  if (vx.V > (1L << 31) - 1) return HSMC_ERR_NOT_SUPPORTED;
  NGrid g;
  int st = build_ngrid(ctx, replica, kDescCut, L_GRID_A, g);
  if (st) return st;
  float *f = (float *)scratch(ctx, L_F, sizeof(float) * (size_t)vx.V);
  if (!f) return HSMC_ERR_NO_MEMORY;
  // the shared score buffer is overwritten: no network's cache stays valid
  ++ctx->score_stamp;
  m->sc_replica = -1;
  Proposal pr;
  st = launch_proposal_scores(ctx, replica, g, vx, m->net, f, &pr);
  if (st) return st;
  hipLaunchKernelGGL(k_mask_blocked, dim3((unsigned)((vx.V + 255) / 256)), dim3(256), 0,
                     ctx->stream, f, pr.flag, vx.V);
  HSMC_HIP(ctx, hipGetLastError());
  if (n_free) *n_free = pr.nfree;
  HSMC_HIP(ctx, hipMemcpyAsync(scores, f, sizeof(float) * vx.V, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HSMC_OK;
}

}  // extern "C"
