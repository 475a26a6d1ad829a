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
// eight of its features. The products are computed transposed (W^T x^T):
// x in this layout is the B operand, and the result returns in the same
// layout (with a fixed feature permutation, the slot order below), so
// message aggregation feeds the update MLP and each product feeds the next
// without a shuffle, an LDS round trip or a barrier. The edge term of every
// message, e_ij We + bm + P_i, is one float32 matrix-core product per edge
// slot as well ([We; bm]^T [e_ij; 1] accumulated onto P_i^T).
//
// The backward pass (training) is hand-written float32 code, one thread per
// vertex, with weight gradients reduced deterministically (block partials in
// block order, then a fixed-order sum).
#include <type_traits>

#include "esr_common.h"

namespace esr {
namespace {

constexpr int H = 32;
constexpr int kB = 256;   // backward / elementwise kernels
// forward kernels: one wavefront (16 vertices) per workgroup (measured
// faster than 2 or 4 wavefronts, or several vertex groups per wavefront)
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
// packed bfloat16 matrices per layer, transposed (output-major), the input
// (K) columns of every H-wide block in slot order (see slot_feature):
// WpT [H][H], WqT [H][H], WuT [H][2H]
constexpr int kPWp = 0, kPWq = H * H, kPWu = 2 * H * H, kPLayer = 4 * H * H;

__device__ __forceinline__ float relu(float x) { return x > 0.0f ? x : 0.0f; }
__device__ __forceinline__ float q16(float x) { return (float)(__bf16)x; }

// Feature slots. Lane l of a wavefront holds vertex l%16 and, in its eight
// slots f, the logical features F(g, f) = 16 (f/4) + 4 g + f%4 with
// g = l/16: the layout in which the matrix cores return W^T x^T. Per-vertex
// activations (h, Q, the saved aggregates a) are stored in slot order,
// position 8 g + f, so a lane moves its features with one 16-byte access.
__host__ __device__ constexpr int slot_feature(int p) {
  return 16 * ((p & 7) >> 2) + 4 * (p >> 3) + (p & 3);
}
__host__ __device__ constexpr int feature_slot(int F) {
  return 8 * ((F & 15) >> 2) + 4 * (F >> 4) + (F & 3);
}
static_assert(feature_slot(slot_feature(13)) == 13 && slot_feature(feature_slot(21)) == 21,
              "slot maps are inverse");
// packed column of logical input k < 2H of the update MLP
__host__ __device__ constexpr int wu_col(int k) { return (k / H) * H + feature_slot(k % H); }

// out (slot layout) = x W for the 16 vertices of the wavefront, computed as
// C^T = W^T x^T: the packed weights WT[m][32 kb + p] = W[32 kb + F(p)][m] are
// the A operand (row m = output feature), x in slot layout is the B operand
// (lane: vertex l%16, K positions 8(l/16)..+7), and C^T comes back with
// lane l holding output features 16 nb + 4(l/16) + v of vertex l%16, which
// is slot 4 nb + v. No LDS round trip, no barrier.
template <int KB>
__device__ __forceinline__ void mfma_slots(const bf16x8 (&x)[KB], const __bf16 *__restrict__ WT,
                                           int ldk, float (&out)[8]) {
  const int l = threadIdx.x & 63, g = l >> 4, j = l & 15;
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    floatx4 c = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const bf16x8 w = *reinterpret_cast<const bf16x8 *>(WT + (16 * nb + j) * ldk + 32 * kb + 8 * g);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, x[kb], c, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) out[4 * nb + v] = c[v];
  }
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
  const Lane ln = lane_of(n);
  const int g = ln.g, j16 = threadIdx.x & 15;
  // h0 = q(relu([u, 1] [W_in; b_in])) as one float32 matrix-core product
  // per 16 output features: lane (g, vertex) supplies component g of
  // u = (r / s, vfeat) (g = 3: the constant 1 that carries b_in), lane
  // (g, m) row g of [W_in; b_in] at feature 16 nb + m; the result is in
  // slot layout.
  float ug = 0.0f;
  if (ln.ok) {
    if (g == 0) {
      const double s = rnorm[0] / sqrt((double)n);
      ug = s > 0.0 ? (float)(r[ln.v] / s) : 0.0f;
    } else if (g < 3) {
      ug = vfeat[2 * ln.v + (g - 1)];
    } else {
      ug = 1.0f;
    }
  }
  float hf[8];
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    const float wa = prm[g * H + 16 * nb + j16];   // rows 0..2: W_in, row 3: b_in
    const floatx4 c = __builtin_amdgcn_mfma_f32_16x16x4f32(wa, ug, floatx4{0.f, 0.f, 0.f, 0.f},
                                                           0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) hf[4 * nb + v] = ln.ok ? q16(relu(c[v])) : 0.0f;
  }
  bf16x8 a[1] = {to_bf16x8(hf)};
  if (ln.ok) *reinterpret_cast<bf16x8 *>(h0 + ln.v * H + 8 * ln.g) = a[0];
  float qv[8];
  mfma_slots<1>(a, WqT0, H, qv);
  if (ln.ok) *reinterpret_cast<bf16x8 *>(Q0 + ln.v * H + 8 * ln.g) = to_bf16x8(qv);
}

// Edges of a vertex, in CSR order. CSR: rowptr / col / efeat (3 floats per
// edge). Slab (SLAB = true): the edges of the wavefront's 16 vertices
// stored slot-major -- edge k of vertex 16 q + j at slot soff[q] + 16 k + j,
// group width (soff[q + 1] - soff[q]) / 16 = its largest degree, ecnt[v] the
// degree, padded slots a valid column and zero features -- so the column and
// feature loads of one slot are one contiguous access per wavefront instead
// of 16 rows' records spread over many cache lines, and they do not wait on
// a row-pointer load. The same edges in the same order either way.
struct EdgeCsr {
  const int32_t *__restrict__ rowptr;
  const int32_t *__restrict__ col;
  const float *__restrict__ feat;
};
struct EdgeSlab {
  const int64_t *__restrict__ soff;
  const uint8_t *__restrict__ cnt;
  const int32_t *__restrict__ col;
  const float *__restrict__ feat;
};

template <bool LAST, typename E>
__global__ __launch_bounds__(kFB) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_gnn_layer(
    int64_t n, const E edges, const __bf16 *__restrict__ h_in,
    const __bf16 *__restrict__ Q_in, const float *__restrict__ lw,
    const __bf16 *__restrict__ lp, const __bf16 *__restrict__ WqT_next,
    __bf16 *__restrict__ h_out, __bf16 *__restrict__ Q_out, float *__restrict__ a_save,
    const float *__restrict__ wout, const double *__restrict__ r,
    const double *__restrict__ d, const double *__restrict__ rnorm, double *__restrict__ z) {
  const Lane ln = lane_of(n);
  const int g = ln.g, j16 = threadIdx.x & 15;
  // edge k of this vertex: position pos(k) in the edge arrays for k < deg;
  // for CSR the row's extent comes first (the edge loads wait on rowptr),
  // for the slab the group's base and width (wave-uniform) and the degree
  int64_t e0 = 0;
  int deg = 0, width = 0;
  bf16x8 hb = bf16x8{};
  if constexpr (std::is_same<E, EdgeSlab>::value) {
    const int64_t q = ln.v >> 4;
    const int64_t b0 = edges.soff[q];
    width = (int)((edges.soff[q + 1] - b0) >> 4);
    e0 = b0 + j16;
    if (ln.ok) deg = edges.cnt[ln.v];
  } else {
    if (ln.ok) {
      e0 = edges.rowptr[ln.v];
      deg = edges.rowptr[ln.v + 1] - (int32_t)e0;
    }
  }
  if (ln.ok) hb = *reinterpret_cast<const bf16x8 *>(h_in + ln.v * H + 8 * g);
  // position of edge k (k clamped to the slots that exist, so every load is
  // valid): CSR e0 + min(k, deg - 1), slab e0 + 16 min(k, width - 1)
  auto pos = [&](int k) -> int64_t {
    if constexpr (std::is_same<E, EdgeSlab>::value)
      return e0 + 16 * (int64_t)(k < width ? k : width - 1);
    else
      return e0 + (k < deg ? k : deg - 1);
  };
  const bool any = std::is_same<E, EdgeSlab>::value ? width > 0 : deg > 0;
  // The first kSlots edges, all loads in flight together. Lane (g, vertex)
  // takes component g of the edge's features (g = 3: the constant 1 that
  // carries bm).
  constexpr int kSlots = 6;
  bf16x8 qs[kSlots];
  float ef[kSlots];
  if (any) {
    int64_t es[kSlots];
    int32_t js[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; ++k) es[k] = pos(k);
#pragma unroll
    for (int k = 0; k < kSlots; ++k) js[k] = edges.col[es[k]];
#pragma unroll
    for (int k = 0; k < kSlots; ++k) {
      qs[k] = *reinterpret_cast<const bf16x8 *>(Q_in + (int64_t)js[k] * H + 8 * g);
      ef[k] = g < 3 ? edges.feat[3 * es[k] + g] : 1.0f;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kSlots; ++k) {
      qs[k] = bf16x8{};
      ef[k] = 0.0f;
    }
  }
  // P = h Wp (overlaps the edge loads)
  float P[8];
  {
    bf16x8 a[1] = {hb};
    mfma_slots<1>(a, lp + kPWp, H, P);
  }
  // Messages. The A operand of the edge product: lane (g, m) holds row g of
  // [We; bm] at output feature 16 nb + m. Every matrix-core instruction runs
  // on the whole wavefront; only the fold is per row, in edge order.
  float wc[2];
#pragma unroll
  for (int nb = 0; nb < 2; ++nb)
    wc[nb] = g < 3 ? lw[Off::We + g * H + 16 * nb + j16] : lw[Off::bm + 16 * nb + j16];
  const floatx4 Pc[2] = {floatx4{P[0], P[1], P[2], P[3]}, floatx4{P[4], P[5], P[6], P[7]}};
  float acc[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) acc[f] = 0.0f;
  auto fold = [&](const floatx4 &o0, const floatx4 &o1, const bf16x8 &qj) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      acc[v] = acc[v] + relu(o0[v] + (float)qj[v]);
      acc[4 + v] = acc[4 + v] + relu(o1[v] + (float)qj[4 + v]);
    }
  };
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const floatx4 o0 = __builtin_amdgcn_mfma_f32_16x16x4f32(wc[0], ef[k], Pc[0], 0, 0, 0);
    const floatx4 o1 = __builtin_amdgcn_mfma_f32_16x16x4f32(wc[1], ef[k], Pc[1], 0, 0, 0);
    if (k < deg) fold(o0, o1, qs[k]);
  }
  // rows longer than kSlots: a trip count uniform over the wavefront (the
  // matrix-core instructions need every lane), loads and fold per row
  int rest = deg - kSlots;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const int o = __shfl_xor(rest, off, 64);
    rest = o > rest ? o : rest;
  }
  for (int t = 0; t < rest; ++t) {
    const bool act = kSlots + t < deg;
    float eft = 0.0f;
    bf16x8 qt = bf16x8{};
    if (act) {
      const int64_t e = std::is_same<E, EdgeSlab>::value ? e0 + 16 * (int64_t)(kSlots + t)
                                                          : e0 + kSlots + t;
      eft = g < 3 ? edges.feat[3 * e + g] : 1.0f;
      qt = *reinterpret_cast<const bf16x8 *>(Q_in + (int64_t)edges.col[e] * H + 8 * g);
    }
    const floatx4 o0 = __builtin_amdgcn_mfma_f32_16x16x4f32(wc[0], eft, Pc[0], 0, 0, 0);
    const floatx4 o1 = __builtin_amdgcn_mfma_f32_16x16x4f32(wc[1], eft, Pc[1], 0, 0, 0);
    if (act) fold(o0, o1, qt);
  }
  if (a_save && ln.ok) {
#pragma unroll
    for (int f = 0; f < 8; ++f) a_save[ln.v * H + 8 * g + f] = acc[f];
  }
  // t = [h, q(a)] Wu + bu ; h' = q(h + relu(t))
  float t[8], hn[8];
  {
    bf16x8 a[2] = {hb, to_bf16x8(acc)};
    mfma_slots<2>(a, lp + kPWu, 2 * H, t);
  }
#pragma unroll
  for (int f = 0; f < 8; ++f)
    hn[f] = q16((float)hb[f] + relu(t[f] + lw[Off::bu + slot_feature(8 * g + f)]));
  const bf16x8 hnb = to_bf16x8(hn);
  if (h_out && ln.ok) *reinterpret_cast<bf16x8 *>(h_out + ln.v * H + 8 * g) = hnb;
  if constexpr (!LAST) {
    bf16x8 a[1] = {hnb};
    float qv[8];
    mfma_slots<1>(a, WqT_next, H, qv);
    if (ln.ok) *reinterpret_cast<bf16x8 *>(Q_out + ln.v * H + 8 * g) = to_bf16x8(qv);
  } else {
    float y = 0.0f;
#pragma unroll
    for (int f = 0; f < 8; ++f) y = y + hn[f] * wout[slot_feature(8 * g + f)];
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
  if (k < kPWq) {            // WpT[n][p] = Wp[F(p)][n]
    const int nn = k / H, kk = slot_feature(k % H);
    v = lw[Off::Wp + kk * H + nn];
  } else if (k < kPWu) {
    const int kq = k - kPWq, nn = kq / H, kk = slot_feature(kq % H);
    v = lw[Off::Wq + kk * H + nn];
  } else {                   // WuT[n][32 b + p] = Wu[32 b + F(p)][n], b < 2
    const int ku = k - kPWu, nn = ku / (2 * H), pk = ku % (2 * H);
    const int kk = (pk / H) * H + slot_feature(pk % H);
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
    x[k] = (float)h[v * H + feature_slot(k)];
    x[H + k] = q16(a[v * H + feature_slot(k)]);
  }
  const __bf16 *WuT = lp + kPWu;
  for (int f = 0; f < H; ++f) {
    float t = 0.0f;
    for (int k = 0; k < 2 * H; ++k) t = t + x[k] * (float)WuT[f * 2 * H + wu_col(k)];
    t = t + lw[Off::bu + f];
    dtv[f] = t > 0.0f ? gh[v * H + f] : 0.0f;
    dt[v * H + f] = dtv[f];
  }
  for (int k = 0; k < 2 * H; ++k) {
    xcat[v * 2 * H + k] = x[k];
    float s = 0.0f;
    for (int f = 0; f < H; ++f) s = s + dtv[f] * (float)WuT[f * 2 * H + wu_col(k)];
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
  for (int k = 0; k < H; ++k) hv[k] = (float)h[v * H + feature_slot(k)];
  const __bf16 *WpT = lp + kPWp;
  for (int f = 0; f < H; ++f) {
    float s = 0.0f;
    for (int k = 0; k < H; ++k) s = s + hv[k] * (float)WpT[f * H + feature_slot(k)];
    P[f] = s;
    dPv[f] = 0.0f;
  }
  for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
    const int64_t j = col[e];
    const float f0 = efeat[3 * (int64_t)e], f1 = efeat[3 * (int64_t)e + 1],
                f2 = efeat[3 * (int64_t)e + 2];
    for (int f = 0; f < H; ++f) {
      // the forward's order: P + e We + bm (matrix core), then + Q_j
      const float pe = (((P[f] + f0 * lw[Off::We + f]) + f1 * lw[Off::We + H + f]) +
                        f2 * lw[Off::We + 2 * H + f]) + lw[Off::bm + f];
      const float pre = pe + (float)Q[j * H + feature_slot(f)];
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
      sp = sp + dP[v * H + f] * (float)WpT[f * H + feature_slot(k)];
      sq = sq + q[f] * (float)WqT[f * H + feature_slot(k)];
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
// bfloat16 X is a stored activation (K = H, columns in slot order).
template <typename TX>
__global__ __launch_bounds__(kB) void k_outer(int64_t m, const TX *__restrict__ X, int K,
                                             const float *__restrict__ D, int N,
                                             float *__restrict__ partial) {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t chunk = (m + nb - 1) / nb;
  const int64_t lo = b * chunk, hi = lo + chunk < m ? lo + chunk : m;
  for (int idx = threadIdx.x; idx < K * N; idx += kB) {
    const int k = idx / N, j = idx % N;
    const int kx = std::is_same<TX, __bf16>::value ? feature_slot(k) : k;
    float s = 0.0f;
    for (int64_t r = lo; r < hi; ++r) {
      const float xv = X ? (float)X[r * K + kx] : 1.0f;
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
namespace {
template <typename E>
int gnn_forward(int L, int64_t n, const E &edges, const float *vfeat, const double *r,
                const double *d, const double *rnorm, const float *params, const void *packed,
                void *h, void *Q, float *a, int save, double *z, hipStream_t s) {
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
      k_gnn_layer<false, E><<<vgrid(n), kFB, 0, s>>>(
          n, edges, hbuf(l), qbuf(l), lw, lp, pk + (int64_t)(l + 1) * kPLayer + kPWq,
          hbuf(l + 1), qbuf(l + 1), asv, nullptr, nullptr, nullptr, nullptr, nullptr);
    } else {
      k_gnn_layer<true, E><<<vgrid(n), kFB, 0, s>>>(
          n, edges, hbuf(l), qbuf(l), lw, lp, nullptr, save ? hbuf(l + 1) : nullptr, nullptr,
          asv, params + o.w_out(), r, d, rnorm, z);
    }
    ESR_LAUNCHED();
  }
  return ESR_OK;
}
} // namespace

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
  return gnn_forward(L, n, EdgeCsr{rowptr, col, efeat}, vfeat, r, d, rnorm, params, packed, h, Q,
                     a, save, z, as_stream(stream));
  ESR_GUARD_END
}

// The same network with the edges in the slab layout (EdgeSlab): soff has
// ceil(n / 16) + 1 entries, ecol and efeat (3 floats) soff[last] slots.
ESR_API int esr_gnn_forward_slab(int L, int64_t n, const int64_t *soff, const uint8_t *ecnt,
                                 const int32_t *ecol, const float *efeat, const float *vfeat,
                                 const double *r, const double *d, const double *rnorm,
                                 const float *params, const void *packed, void *h, void *Q,
                                 float *a, int save, double *z, void *stream) {
  ESR_GUARD_BEGIN
  ESR_REQUIRE(L >= 1 && n >= 1 && soff && ecnt && ecol && efeat && vfeat && r && d && rnorm &&
                  params && packed && h && Q && z,
              "esr_gnn_forward_slab: null argument");
  ESR_REQUIRE(!save || a, "esr_gnn_forward_slab: save needs the a buffers");
  return gnn_forward(L, n, EdgeSlab{soff, ecnt, ecol, efeat}, vfeat, r, d, rnorm, params, packed,
                     h, Q, a, save, z, as_stream(stream));
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
