// Learned preconditioner: an edge message-passing network z = M(r) on the
// mesh graph (no reference counterpart; the reference's preconditioner slot,
// easier/numeric/preconditioner/__init__.py:4, is empty).
//
//   s = sqrt(mean r^2);  u_i = (r_i/s, vfeat_i)
//   h0_i = q(relu(u_i W_in + b_in))
//   layer l: P = h Wp; Q = q(h Wq); a_i = sum_j relu(P_i + Q_j + e_ij We + bm)
//            h <- q(h + relu([h, q(a)] Wu + bu))
//   z_i = (r_i + s (h_i w_out + b_out)) / d_i
// q() = round to bfloat16. The per-vertex products (h Wp, h Wq, [h,a] Wu) run
// on the matrix cores: v_mfma_f32_16x16x32_bf16, bfloat16 operands, float32
// accumulation. A wavefront owns 16 vertices; lane l holds vertex l%16 and
// features 8*(l/16)..+7, which is the MFMA A-operand layout, so message
// aggregation feeds the update MLP without a shuffle. MFMA results come back
// row-per-lane through LDS.
//
// The backward pass (training) is hand-written float32 code, one thread per
// vertex, with weight gradients reduced deterministically (block partials in
// block order, then a fixed-order sum).
#include "esr_common.h"

namespace esr {
namespace {

constexpr int H = 32;
constexpr int kB = 256;   // backward / elementwise kernels
// forward kernels: one wavefront per workgroup, so the LDS round trips of
// the MFMA layout changes (c_to_rows) synchronise one wave, not four
constexpr int kFB = 64;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

// flat float32 parameter layout (must match numeric/preconditioner/learned.py)
struct Off {
  int L;
  __host__ __device__ int W_in() const { return 0; }
  __host__ __device__ int b_in() const { return 3 * H; }
  __host__ __device__ int layer(int l) const { return 4 * H + l * kLayer; }
  static constexpr int Wp = 0, Wq = H * H, We = 2 * H * H, bm = 2 * H * H + 3 * H,
                       Wu = 2 * H * H + 4 * H, bu = 4 * H * H + 4 * H,
                       kLayer = 4 * H * H + 5 * H;
  __host__ __device__ int w_out() const { return 4 * H + L * kLayer; }
  __host__ __device__ int b_out() const { return w_out() + H; }
  __host__ __device__ int total() const { return b_out() + 1; }
};
// packed bfloat16 matrices per layer, transposed (output-major):
// WpT [H][H], WqT [H][H], WuT [H][2H]
constexpr int kPWp = 0, kPWq = H * H, kPWu = 2 * H * H, kPLayer = 4 * H * H;

__device__ __forceinline__ float relu(float x) { return x > 0.0f ? x : 0.0f; }
__device__ __forceinline__ float q16(float x) { return (float)(__bf16)x; }

// C[16][32] (two 16x16 blocks) += A[16][32*KB] * W[32*KB][32], W given
// transposed (WT[n][k], row length ldk). Lane layout: A chunk kb holds
// A[l%16][32kb + 8(l/16) .. +7]; c[nb][v] = C[4(l/16)+v][16nb + l%16].
template <int KB>
__device__ __forceinline__ void mfma_rows(const bf16x8 (&a)[KB], const __bf16 *__restrict__ WT,
                                          int ldk, floatx4 (&c)[2]) {
  const int l = threadIdx.x & 63, g = l >> 4, j = l & 15;
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    c[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const bf16x8 b = *reinterpret_cast<const bf16x8 *>(WT + (16 * nb + j) * ldk + 32 * kb + 8 * g);
      c[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kb], b, c[nb], 0, 0, 0);
    }
  }
}

// MFMA C layout -> row layout (lane l: row l%16, columns 8(l/16)..+7).
// Every wave of the block must call it (block barriers).
__device__ __forceinline__ void c_to_rows(const floatx4 (&c)[2], float (*T)[33], float (&out)[8]) {
  const int l = threadIdx.x & 63, g = l >> 4, j = l & 15;
#pragma unroll
  for (int nb = 0; nb < 2; ++nb)
#pragma unroll
    for (int v = 0; v < 4; ++v) T[4 * g + v][16 * nb + j] = c[nb][v];
  __syncthreads();
#pragma unroll
  for (int f = 0; f < 8; ++f) out[f] = T[j][8 * g + f];
  __syncthreads();
}

__device__ __forceinline__ bf16x8 to_bf16x8(const float (&x)[8]) {
  bf16x8 r;
#pragma unroll
  for (int f = 0; f < 8; ++f) r[f] = (__bf16)x[f];
  return r;
}

struct Lane {
  int64_t v; // vertex
  int g;     // feature group
  bool ok;
};
__device__ __forceinline__ Lane lane_of(int64_t n) {
  const int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  Lane ln;
  ln.v = (blk * (kFB / 64) + w) * 16 + (l & 15);
  ln.g = l >> 4;
  ln.ok = ln.v < n;
  return ln;
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kFB) void k_gnn_encode(int64_t n, const double *__restrict__ r,
                                                  const float *__restrict__ vfeat,
                                                  const double *__restrict__ rnorm,
                                                  const float *__restrict__ prm,
                                                  const __bf16 *__restrict__ WqT0,
                                                  __bf16 *__restrict__ h0, __bf16 *__restrict__ Q0) {
  __shared__ float T[kFB / 64][16][33];
  const Lane ln = lane_of(n);
  const double s = rnorm[0] / sqrt((double)n);
  float u0 = 0.f, u1 = 0.f, u2 = 0.f;
  if (ln.ok) {
    u0 = s > 0.0 ? (float)(r[ln.v] / s) : 0.0f;
    u1 = vfeat[2 * ln.v];
    u2 = vfeat[2 * ln.v + 1];
  }
  float hf[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) {
    const int F = 8 * ln.g + f;
    const float x = ((u0 * prm[F] + u1 * prm[H + F]) + u2 * prm[2 * H + F]) + prm[3 * H + F];
    hf[f] = ln.ok ? q16(relu(x)) : 0.0f;
  }
  bf16x8 a[1] = {to_bf16x8(hf)};
  if (ln.ok) *reinterpret_cast<bf16x8 *>(h0 + ln.v * H + 8 * ln.g) = a[0];
  floatx4 c[2];
  mfma_rows<1>(a, WqT0, H, c);
  float qv[8];
  c_to_rows(c, T[threadIdx.x >> 6], qv);
  if (ln.ok) *reinterpret_cast<bf16x8 *>(Q0 + ln.v * H + 8 * ln.g) = to_bf16x8(qv);
}

template <bool LAST>
__global__ __launch_bounds__(kFB) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_gnn_layer(
    int64_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const float *__restrict__ efeat, const __bf16 *__restrict__ h_in,
    const __bf16 *__restrict__ Q_in, const float *__restrict__ lw,
    const __bf16 *__restrict__ lp, const __bf16 *__restrict__ WqT_next,
    __bf16 *__restrict__ h_out, __bf16 *__restrict__ Q_out, float *__restrict__ a_save,
    const float *__restrict__ wout, const double *__restrict__ r,
    const double *__restrict__ d, const double *__restrict__ rnorm, double *__restrict__ z) {
  __shared__ float T[kFB / 64][16][33];
  float(*Tw)[33] = T[threadIdx.x >> 6];
  const Lane ln = lane_of(n);
  const int g = ln.g;
  bf16x8 hb = bf16x8{};
  if (ln.ok) hb = *reinterpret_cast<const bf16x8 *>(h_in + ln.v * H + 8 * g);
  // P = h Wp
  floatx4 c[2];
  {
    bf16x8 a[1] = {hb};
    mfma_rows<1>(a, lp + kPWp, H, c);
  }
  float P[8];
  c_to_rows(c, Tw, P);
  // messages
  float we0[8], we1[8], we2[8], bm[8], acc[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) {
    const int F = 8 * g + f;
    we0[f] = lw[Off::We + F];
    we1[f] = lw[Off::We + H + F];
    we2[f] = lw[Off::We + 2 * H + F];
    bm[f] = lw[Off::bm + F];
    acc[f] = 0.0f;
  }
  if (ln.ok) {
    const int32_t e0 = rowptr[ln.v], e1 = rowptr[ln.v + 1];
    auto message = [&](const bf16x8 &qj, float f0, float f1, float f2) {
#pragma unroll
      for (int f = 0; f < 8; ++f) {
        const float ew = (f0 * we0[f] + f1 * we1[f]) + f2 * we2[f];
        const float pre = ((P[f] + (float)qj[f]) + ew) + bm[f];
        acc[f] = acc[f] + relu(pre);
      }
    };
    if (e1 > e0) {
      // the first kSlots edges: positions clamped to the row, so every load
      // is valid and all are in flight together; folded in edge order
      constexpr int kSlots = 6;
      int32_t es[kSlots], js[kSlots];
#pragma unroll
      for (int k = 0; k < kSlots; ++k) es[k] = e0 + k < e1 ? e0 + k : e1 - 1;
#pragma unroll
      for (int k = 0; k < kSlots; ++k) js[k] = col[es[k]];
      bf16x8 qs[kSlots];
      float fs[kSlots][3];
#pragma unroll
      for (int k = 0; k < kSlots; ++k) {
        qs[k] = *reinterpret_cast<const bf16x8 *>(Q_in + (int64_t)js[k] * H + 8 * g);
        fs[k][0] = efeat[3 * (int64_t)es[k]];
        fs[k][1] = efeat[3 * (int64_t)es[k] + 1];
        fs[k][2] = efeat[3 * (int64_t)es[k] + 2];
      }
#pragma unroll
      for (int k = 0; k < kSlots; ++k)
        if (e0 + k < e1) message(qs[k], fs[k][0], fs[k][1], fs[k][2]);
      for (int32_t e = e0 + kSlots; e < e1; ++e) {
        const int64_t j = col[e];
        message(*reinterpret_cast<const bf16x8 *>(Q_in + j * H + 8 * g), efeat[3 * (int64_t)e],
                efeat[3 * (int64_t)e + 1], efeat[3 * (int64_t)e + 2]);
      }
    }
    if (a_save) {
#pragma unroll
      for (int f = 0; f < 8; ++f) a_save[ln.v * H + 8 * g + f] = acc[f];
    }
  }
  // t = [h, q(a)] Wu + bu ; h' = q(h + relu(t))
  {
    bf16x8 a[2] = {hb, to_bf16x8(acc)};
    mfma_rows<2>(a, lp + kPWu, 2 * H, c);
  }
  float t[8], hn[8];
  c_to_rows(c, Tw, t);
#pragma unroll
  for (int f = 0; f < 8; ++f) hn[f] = q16((float)hb[f] + relu(t[f] + lw[Off::bu + 8 * g + f]));
  const bf16x8 hnb = to_bf16x8(hn);
  if (h_out && ln.ok) *reinterpret_cast<bf16x8 *>(h_out + ln.v * H + 8 * g) = hnb;
  if constexpr (!LAST) {
    bf16x8 a[1] = {hnb};
    mfma_rows<1>(a, WqT_next, H, c);
    float qv[8];
    c_to_rows(c, Tw, qv);
    if (ln.ok) *reinterpret_cast<bf16x8 *>(Q_out + ln.v * H + 8 * g) = to_bf16x8(qv);
  } else {
    float y = 0.0f;
#pragma unroll
    for (int f = 0; f < 8; ++f) y = y + hn[f] * wout[8 * g + f];
    // the four lanes of a vertex: l, l+16, l+32, l+48
    const int l = threadIdx.x & 63;
    const float y1 = __shfl(y, (l & 15) + 16, 64);
    const float y2 = __shfl(y, (l & 15) + 32, 64);
    const float y3 = __shfl(y, (l & 15) + 48, 64);
    if (g == 0 && ln.ok) {
      const float yt = ((y + y1) + y2) + y3 + wout[H];
      const double s = rnorm[0] / sqrt((double)n);
      z[ln.v] = (r[ln.v] + s * (double)yt) / d[ln.v];
    }
  }
}

__global__ void k_gnn_pack(int L, const float *__restrict__ prm, __bf16 *__restrict__ packed) {
  const Off o{L};
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = kPLayer;
  if (idx >= L * per) return;
  const int l = idx / per, k = idx % per;
  const float *lw = prm + o.layer(l);
  float v;
  if (k < kPWq) {            // WpT[n][k'] = Wp[k'][n]
    const int nn = k / H, kk = k % H;
    v = lw[Off::Wp + kk * H + nn];
  } else if (k < kPWu) {
    const int kq = k - kPWq, nn = kq / H, kk = kq % H;
    v = lw[Off::Wq + kk * H + nn];
  } else {                   // WuT[n][k'] = Wu[k'][n], k' < 2H
    const int ku = k - kPWu, nn = ku / (2 * H), kk = ku % (2 * H);
    v = lw[Off::Wu + kk * H + nn];
  }
  packed[idx] = (__bf16)v;
}

// ---------------------------------------------------------------------------
// backward (float32, one thread per vertex / edge owner)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t tid_global() {
  return xcd_remap(blockIdx.x, gridDim.x) * kB + threadIdx.x;
}

// dy = dz s / d; gh = dy w_out
__global__ __launch_bounds__(kB) void k_gnn_bwd_out(int64_t n, const double *__restrict__ dz,
                                                   const double *__restrict__ d,
                                                   const double *__restrict__ rnorm,
                                                   const float *__restrict__ wout,
                                                   float *__restrict__ dy, float *__restrict__ gh) {
  const int64_t v = tid_global();
  if (v >= n) return;
  const double s = rnorm[0] / sqrt((double)n);
  const float y = (float)(dz[v] * s / d[v]);
  dy[v] = y;
  for (int f = 0; f < H; ++f) gh[v * H + f] = y * wout[f];
}

// update MLP of layer l: x = [h, q(a)], t = x Wu + bu, dt = gh * [t > 0];
// writes x (for dWu), dt, gtot = gh + (dt Wu^T)[:H], da = (dt Wu^T)[H:].
__global__ __launch_bounds__(kB) void k_gnn_bwd_upd(int64_t n, const __bf16 *__restrict__ h,
                                                   const float *__restrict__ a,
                                                   const float *__restrict__ lw,
                                                   const __bf16 *__restrict__ lp,
                                                   const float *__restrict__ gh,
                                                   float *__restrict__ xcat, float *__restrict__ dt,
                                                   float *__restrict__ gtot, float *__restrict__ da) {
  const int64_t v = tid_global();
  if (v >= n) return;
  float x[2 * H], dtv[H];
  for (int k = 0; k < H; ++k) {
    x[k] = (float)h[v * H + k];
    x[H + k] = q16(a[v * H + k]);
  }
  const __bf16 *WuT = lp + kPWu;
  for (int f = 0; f < H; ++f) {
    float t = 0.0f;
    for (int k = 0; k < 2 * H; ++k) t = t + x[k] * (float)WuT[f * 2 * H + k];
    t = t + lw[Off::bu + f];
    dtv[f] = t > 0.0f ? gh[v * H + f] : 0.0f;
    dt[v * H + f] = dtv[f];
  }
  for (int k = 0; k < 2 * H; ++k) {
    xcat[v * 2 * H + k] = x[k];
    float s = 0.0f;
    for (int f = 0; f < H; ++f) s = s + dtv[f] * (float)WuT[f * 2 * H + k];
    if (k < H) gtot[v * H + k] = gh[v * H + k] + s;
    else da[v * H + (k - H)] = s;
  }
}

// messages of layer l: recompute pre_ij, dpre = da_i [pre > 0] (stored per
// edge), dP_i = sum_j dpre.
__global__ __launch_bounds__(kB) void k_gnn_bwd_edge(int64_t n, const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ col,
                                                    const float *__restrict__ efeat,
                                                    const __bf16 *__restrict__ h,
                                                    const __bf16 *__restrict__ Q,
                                                    const float *__restrict__ lw,
                                                    const __bf16 *__restrict__ lp,
                                                    const float *__restrict__ da,
                                                    float *__restrict__ dpre, float *__restrict__ dP) {
  const int64_t v = tid_global();
  if (v >= n) return;
  float hv[H], P[H], dPv[H];
  for (int k = 0; k < H; ++k) hv[k] = (float)h[v * H + k];
  const __bf16 *WpT = lp + kPWp;
  for (int f = 0; f < H; ++f) {
    float s = 0.0f;
    for (int k = 0; k < H; ++k) s = s + hv[k] * (float)WpT[f * H + k];
    P[f] = s;
    dPv[f] = 0.0f;
  }
  for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
    const int64_t j = col[e];
    const float f0 = efeat[3 * (int64_t)e], f1 = efeat[3 * (int64_t)e + 1],
                f2 = efeat[3 * (int64_t)e + 2];
    for (int f = 0; f < H; ++f) {
      const float ew = (f0 * lw[Off::We + f] + f1 * lw[Off::We + H + f]) + f2 * lw[Off::We + 2 * H + f];
      const float pre = ((P[f] + (float)Q[j * H + f]) + ew) + lw[Off::bm + f];
      const float g = pre > 0.0f ? da[v * H + f] : 0.0f;
      dpre[(int64_t)e * H + f] = g;
      dPv[f] = dPv[f] + g;
    }
  }
  for (int f = 0; f < H; ++f) dP[v * H + f] = dPv[f];
}

// dQ_j = sum over edges (i <- j) of dpre (the reverse edge of each edge of
// row j); gh = gtot + dP Wp^T + dQ Wq^T (input gradient of layer l).
__global__ __launch_bounds__(kB) void k_gnn_bwd_node(int64_t n, const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ rev,
                                                    const float *__restrict__ dpre,
                                                    const float *__restrict__ dP,
                                                    const float *__restrict__ gtot,
                                                    const __bf16 *__restrict__ lp,
                                                    float *__restrict__ dQ, float *__restrict__ gh) {
  const int64_t v = tid_global();
  if (v >= n) return;
  float q[H];
  for (int f = 0; f < H; ++f) q[f] = 0.0f;
  for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
    const int64_t re = rev[e];
    for (int f = 0; f < H; ++f) q[f] = q[f] + dpre[re * H + f];
  }
  const __bf16 *WpT = lp + kPWp, *WqT = lp + kPWq;
  for (int f = 0; f < H; ++f) dQ[v * H + f] = q[f];
  for (int k = 0; k < H; ++k) {
    float s = gtot[v * H + k];
    float sp = 0.0f, sq = 0.0f;
    for (int f = 0; f < H; ++f) {
      sp = sp + dP[v * H + f] * (float)WpT[f * H + k];
      sq = sq + q[f] * (float)WqT[f * H + k];
    }
    gh[v * H + k] = (s + sp) + sq;
  }
}

// encoder: pre0 = u W_in + b_in; dpre0 = gh [pre0 > 0]; writes u and dpre0.
__global__ __launch_bounds__(kB) void k_gnn_bwd_enc(int64_t n, const double *__restrict__ r,
                                                   const float *__restrict__ vfeat,
                                                   const double *__restrict__ rnorm,
                                                   const float *__restrict__ prm,
                                                   const float *__restrict__ gh,
                                                   float *__restrict__ u, float *__restrict__ dpre) {
  const int64_t v = tid_global();
  if (v >= n) return;
  const double s = rnorm[0] / sqrt((double)n);
  const float u0 = s > 0.0 ? (float)(r[v] / s) : 0.0f, u1 = vfeat[2 * v], u2 = vfeat[2 * v + 1];
  u[3 * v] = u0;
  u[3 * v + 1] = u1;
  u[3 * v + 2] = u2;
  for (int f = 0; f < H; ++f) {
    const float x = ((u0 * prm[f] + u1 * prm[H + f]) + u2 * prm[2 * H + f]) + prm[3 * H + f];
    dpre[v * H + f] = x > 0.0f ? gh[v * H + f] : 0.0f;
  }
}

// partial[b][k*N + j] = sum over rows of block b of X[r][k] * D[r][j]
// (X == nullptr: X = 1, K = 1); rows split into nb equal chunks in order.
template <typename TX>
__global__ __launch_bounds__(kB) void k_outer(int64_t m, const TX *__restrict__ X, int K,
                                             const float *__restrict__ D, int N,
                                             float *__restrict__ partial) {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t chunk = (m + nb - 1) / nb;
  const int64_t lo = b * chunk, hi = lo + chunk < m ? lo + chunk : m;
  for (int idx = threadIdx.x; idx < K * N; idx += kB) {
    const int k = idx / N, j = idx % N;
    float s = 0.0f;
    for (int64_t r = lo; r < hi; ++r) {
      const float xv = X ? (float)X[r * K + k] : 1.0f;
      s = s + xv * D[r * N + j];
    }
    partial[b * (int64_t)K * N + idx] = s;
  }
}

// grad[off + idx] (+)= sum_b partial[b][idx], b in order
__global__ __launch_bounds__(kB) void k_partial_sum(int64_t nb, int size,
                                                   const float *__restrict__ partial,
                                                   float *__restrict__ grad, int accumulate) {
  const int idx = blockIdx.x * kB + threadIdx.x;
  if (idx >= size) return;
  float s = 0.0f;
  for (int64_t b = 0; b < nb; ++b) s = s + partial[b * size + idx];
  grad[idx] = accumulate ? grad[idx] + s : s;
}

__global__ __launch_bounds__(kB) void k_adam(int64_t np_, float *__restrict__ p,
                                            const float *__restrict__ g, float *__restrict__ m,
                                            float *__restrict__ v, float lr, float b1, float b2,
                                            float eps, float c1, float c2, float clip) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i >= np_) return;
  float gi = g[i];
  if (clip > 0.0f) gi = gi > clip ? clip : (gi < -clip ? -clip : gi);
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] = p[i] - lr * (mi / c1) / (cr_sqrt(vi / c2) + eps);
}

} // namespace
} // namespace esr

using namespace esr;

static unsigned vgrid(int64_t n) { return (unsigned)ceil_div(n, 16 * (kFB / 64)); } // 16 vertices / wave
static unsigned tgrid(int64_t n) { return (unsigned)ceil_div(n, kB); }

ESR_API int esr_gnn_sizes(int L, int64_t *n_params, int64_t *n_packed) {
  ESR_REQUIRE(L >= 1 && L <= 16 && n_params && n_packed, "esr_gnn_sizes: bad args");
  *n_params = Off{L}.total();
  *n_packed = (int64_t)L * kPLayer;
  return ESR_OK;
}

ESR_API int esr_gnn_pack(int L, const float *params, void *packed, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(L >= 1 && params && packed, "esr_gnn_pack: bad args");
  const int tot = L * kPLayer;
  k_gnn_pack<<<(unsigned)ceil_div(tot, kB), kB, 0, as_stream(stream)>>>(L, params, (__bf16 *)packed);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}

// z = M(r). h: (save ? L+1 : 2) buffers of n*H bf16; Q: (save ? L : 2)
// buffers; a: L buffers of n*H float32 when save != 0 (else NULL).
// rnorm: device float64 scalar ||r||.
ESR_API int esr_gnn_forward(int L, int64_t n, const int32_t *rowptr, const int32_t *col,
                            const float *efeat, const float *vfeat, const double *r,
                            const double *d, const double *rnorm, const float *params,
                            const void *packed, void *h, void *Q, float *a, int save,
                            double *z, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(L >= 1 && n >= 1 && rowptr && col && efeat && vfeat && r && d && rnorm && params &&
                  packed && h && Q && z,
              "esr_gnn_forward: null argument");
  ESR_REQUIRE(!save || a, "esr_gnn_forward: save needs the a buffers");
  hipStream_t s = as_stream(stream);
  const Off o{L};
  const __bf16 *pk = (const __bf16 *)packed;
  __bf16 *hb = (__bf16 *)h, *qb = (__bf16 *)Q;
  auto hbuf = [&](int l) { return hb + (int64_t)(save ? l : (l & 1)) * n * H; };
  auto qbuf = [&](int l) { return qb + (int64_t)(save ? l : (l & 1)) * n * H; };
  k_gnn_encode<<<vgrid(n), kFB, 0, s>>>(n, r, vfeat, rnorm, params, pk + kPWq, hbuf(0), qbuf(0));
  ESR_LAUNCHED();
  for (int l = 0; l < L; ++l) {
    const float *lw = params + o.layer(l);
    const __bf16 *lp = pk + (int64_t)l * kPLayer;
    float *asv = save ? a + (int64_t)l * n * H : nullptr;
    if (l + 1 < L) {
      k_gnn_layer<false><<<vgrid(n), kFB, 0, s>>>(
          n, rowptr, col, efeat, hbuf(l), qbuf(l), lw, lp, pk + (int64_t)(l + 1) * kPLayer + kPWq,
          hbuf(l + 1), qbuf(l + 1), asv, nullptr, nullptr, nullptr, nullptr, nullptr);
    } else {
      k_gnn_layer<true><<<vgrid(n), kFB, 0, s>>>(
          n, rowptr, col, efeat, hbuf(l), qbuf(l), lw, lp, nullptr, save ? hbuf(l + 1) : nullptr,
          nullptr, asv, params + o.w_out(), r, d, rnorm, z);
    }
    ESR_LAUNCHED();
  }
  return ESR_OK;
  ESR_GUARD_END
}

// Gradient of sum_i dz_i z_i w.r.t. the parameters, written into grad (same
// layout as params). Needs the saved forward (save = 1) buffers h, Q, a, the
// reverse-edge index rev, and scratch: work (float32) of
// n*(3H + 2H + H + H + H + 3) + nnz*H + nb*2H*H floats, nb = 256.
ESR_API int esr_gnn_backward(int L, int64_t n, int64_t nnz, const int32_t *rowptr,
                             const int32_t *col, const int32_t *rev, const float *efeat,
                             const float *vfeat, const double *r, const double *d,
                             const double *rnorm, const float *params, const void *packed,
                             const void *h, const void *Q, const float *a, const double *dz,
                             float *grad, float *work, int64_t work_floats, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(L >= 1 && n >= 1 && nnz >= 1, "esr_gnn_backward: bad sizes");
  const int64_t nb = 256;
  const int64_t need = n * (2 * H + 6 * H + 3) + nnz * H + nb * 2 * H * H;
  ESR_REQUIRE(work && work_floats >= need, "esr_gnn_backward: work needs %lld floats",
              (long long)need);
  hipStream_t s = as_stream(stream);
  const Off o{L};
  const __bf16 *pk = (const __bf16 *)packed;
  const __bf16 *hb = (const __bf16 *)h, *qb = (const __bf16 *)Q;
  float *xcat = work;                 // n*2H
  float *dt = xcat + n * 2 * H;       // n*H
  float *gtot = dt + n * H;           // n*H
  float *da = gtot + n * H;           // n*H
  float *dP = da + n * H;             // n*H
  float *gh = dP + n * H;             // n*H
  float *dQ = gh + n * H;             // n*H
  float *u = dQ + n * H;              // n*3
  float *dpre = u + n * 3;            // nnz*H
  float *partial = dpre + nnz * H;    // nb*2H*H
  float *dy = da;                     // reuse (only before the first layer)
  auto reduce = [&](auto X, int K, const float *D, int N, int64_t m, int off) -> int {
    k_outer<<<(unsigned)nb, kB, 0, s>>>(m, X, K, D, N, partial);
    k_partial_sum<<<(unsigned)ceil_div(K * N, kB), kB, 0, s>>>(nb, K * N, partial, grad + off, 0);
    ESR_LAUNCHED();
    return ESR_OK;
  };
  k_gnn_bwd_out<<<tgrid(n), kB, 0, s>>>(n, dz, d, rnorm, params + o.w_out(), dy, gh);
  ESR_LAUNCHED();
  reduce(hb + (int64_t)L * n * H, H, dy, 1, n, o.w_out());
  reduce((const float *)nullptr, 1, dy, 1, n, o.b_out());
  for (int l = L - 1; l >= 0; --l) {
    const float *lw = params + o.layer(l);
    const __bf16 *lp = pk + (int64_t)l * kPLayer;
    const __bf16 *hl = hb + (int64_t)l * n * H, *ql = qb + (int64_t)l * n * H;
    k_gnn_bwd_upd<<<tgrid(n), kB, 0, s>>>(n, hl, a + (int64_t)l * n * H, lw, lp, gh, xcat, dt,
                                          gtot, da);
    ESR_LAUNCHED();
    reduce(xcat, 2 * H, dt, H, n, o.layer(l) + Off::Wu);
    reduce((const float *)nullptr, 1, dt, H, n, o.layer(l) + Off::bu);
    k_gnn_bwd_edge<<<tgrid(n), kB, 0, s>>>(n, rowptr, col, efeat, hl, ql, lw, lp, da, dpre, dP);
    ESR_LAUNCHED();
    reduce(efeat, 3, dpre, H, nnz, o.layer(l) + Off::We);
    reduce((const float *)nullptr, 1, dpre, H, nnz, o.layer(l) + Off::bm);
    k_gnn_bwd_node<<<tgrid(n), kB, 0, s>>>(n, rowptr, rev, dpre, dP, gtot, lp, dQ, gh);
    ESR_LAUNCHED();
    reduce(hl, H, dP, H, n, o.layer(l) + Off::Wp);
    reduce(hl, H, dQ, H, n, o.layer(l) + Off::Wq);
  }
  k_gnn_bwd_enc<<<tgrid(n), kB, 0, s>>>(n, r, vfeat, rnorm, params, gh, u, dpre);
  ESR_LAUNCHED();
  reduce((const float *)u, 3, dpre, H, n, o.W_in());
  reduce((const float *)nullptr, 1, dpre, H, n, o.b_in());
  return ESR_OK;
  ESR_GUARD_END
}

ESR_API int esr_gnn_adam(int64_t np_, float *params, const float *grad, float *m, float *v,
                         double lr, double b1, double b2, double eps, int64_t step, double clip,
                         void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(np_ >= 1 && params && grad && m && v && step >= 1, "esr_gnn_adam: bad args");
  const double c1 = 1.0 - pow(b1, (double)step), c2 = 1.0 - pow(b2, (double)step);
  k_adam<<<(unsigned)ceil_div(np_, kB), kB, 0, as_stream(stream)>>>(
      np_, params, grad, m, v, (float)lr, (float)b1, (float)b2, (float)eps, (float)c1, (float)c2,
      (float)clip);
  ESR_LAUNCHED();
  return ESR_OK;
  ESR_GUARD_END
}
