// Differentiable layered offset / normalized min-sum decoder: forward and adjoint.
//
// No counterpart in the reference.  The forward is the reference's float min-sum
// layer update (src/LDPCDec/decode_functions.cpp:11-103: A_j = max(|q_j| - beta, 0),
// exclusive minimum capped at max_row_metric, sign of the product of the other
// signs) run for a fixed T iterations (no syndrome exit) with an offset beta[t][l]
// and a scale alpha[t][l] per (iteration, layer):  M_j = alpha * min(cap, min_{k!=j} A_k).
// With alpha = 1 and a constant beta it equals the reference decoder on codewords
// that need all T iterations.
//
// Memory: the forward keeps checkpoints only at the starts of NS = ceil(T / K) segments
// (posteriors P and C2V C; K is chosen by the trainer to minimise the memory, of order
// sqrt(T)), and one working C2V array.  The segments are [0, f), [f, f + K), ...,
// [T - K, T) with f = T - (NS - 1) K: the last one, whose V2C messages the forward keeps,
// is a full K long, so the adjoint runs T - K iterations again (not T - f).  The adjoint takes the segments in reverse: it restarts
// from the segment's checkpoint and runs the segment's iterations again, keeping their
// V2C messages q (the only values the reverse sweep needs), then sweeps those iterations
// backwards, layer by layer:
//   g_out_j = gP[v_j] + gC[e_j]           (C2V_j feeds the posterior and iteration t+1)
//   g_m_j   = sign_j * g_out_j,  g_alpha += g_m_j * min(cap, x_j),  g_A[src_j] += alpha * g_m_j (x_j < cap)
//   g_q_j   = gP[v_j] + [|q_j| > beta] * sign(q_j) * g_A_j,  g_beta -= [|q_j| > beta] * g_A_j
//             (g_A_j is nonzero only at the argmin and the second minimum)
//   gP[v_j] <- g_q_j,  gC[e_j] <- -g_q_j    (gradient w.r.t. the C2V entering iteration t)
// where x_j is the exclusive minimum and src_j the edge it came from: the row's argmin
// for every other edge, its second minimum for the argmin itself.  Per codeword and
// iteration this stores (N + E)/K + E values instead of N + E, and recomputes each
// iteration once.  Parameter gradients are written per codeword and summed by
// k_train_sum (deterministic).
//
// One wave per codeword, lane owns rows lane, lane+64, ... of every layer.  Every block
// first copies the code's edge words and layer weights into the front of its LDS.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "decode_core.h"
#include "internal.h"
#include "train_core.h"

namespace fecsim {

namespace {

using namespace train_core;

constexpr int kTrainMaxW = 32;


template <typename F>
struct TrainView {
  F* P;    // posteriors (LDS or global)
  F* gP;   // their adjoints (LDS or global)
};

template <typename F, bool GP>
__device__ __forceinline__ TrainView<F> train_view(const TrainArgs& a, uint64_t cw, char* smem, int wave) {
  TrainView<F> v;
  if constexpr (GP) {
    v.P = reinterpret_cast<F*>(a.gscratch) + cw * (uint64_t)(2 * a.n);
    v.gP = v.P + a.n;
  } else {
    v.P = reinterpret_cast<F*>(smem + a.tab_bytes) + (size_t)wave * 2 * a.n_pad;
    v.gP = v.P + a.n_pad;
  }
  return v;
}


// one check row of the forward: C_old -> C_new, P updated in place; optionally keeps q
// (with QT: the V2C value before its saturation). MAXW bounds the row weight w (the
// unrolled loops stop there).
template <typename F, int MAXW, bool QT>
__device__ __forceinline__ void fwd_row(const TrainArgs& a, const uint32_t* __restrict__ edges, F* P, const F* Cold,
                                        F* Cnew, F* Qs, int e0, int w, int r, F al, F be, F cap) {
  const int Z = a.z;
  F q[MAXW];
  int vi[MAXW];
  F m1 = F(INFINITY), m2 = F(INFINITY);
  int im = -1, sg = 0;
#pragma unroll
  for (int j = 0; j < MAXW; ++j) {
    if (j < w) {
      vi[j] = edge_var(edges[e0 + j], r, Z);
      const int ce = (e0 + j) * Z + r;
      const F xq = P[vi[j]] - Cold[ce];
      if (Qs) Qs[ce] = xq;
      q[j] = QT ? sat8f(xq) : xq;
      const F A = mag_a<F, QT>(q[j], be);
      sg ^= sbit(q[j]);
      if (A < m1) { m2 = m1; m1 = A; im = j; }
      else if (A < m2) m2 = A;
    }
  }
#pragma unroll
  for (int j = 0; j < MAXW; ++j) {
    if (j < w) {
      const F x = (j == im) ? m2 : m1;
      const F m = al * (cap < x ? cap : x);
      const bool neg = sg ^ sbit(q[j]);
      const F rn = neg ? -m : m;
      P[vi[j]] = QT ? sat8f(q[j] + rn) : q[j] + rn;
      if (Cnew) Cnew[(e0 + j) * Z + r] = rn;
    }
  }
}

// one check row of the reverse sweep (see the head of this file); sets the row's
// parameter-gradient terms *ga, *gb
// with QT a saturated value passes no gradient (the clamps), and the offset's rounding
// passes it straight through (d round(b) / db taken as 1)
//
// Summation order (train_reg.hip follows it operation for operation): with gm_j the C2V
// adjoint times the C2V sign, S = sum over the edges other than the argmin in edge order,
// the scale's term is c1 S + c2 gm_argmin, A_argmin's gradient is alpha S, the second
// minimum's alpha gm_argmin; only the argmin and the second-minimum edge take a term back
// (each if its magnitude was positive), and the offset's term is minus the sum of those
// two.
template <typename F, int MAXW, bool QT>
__device__ __forceinline__ void rev_row(const uint32_t* __restrict__ edges, F* gP, F* gC, const F* Qs, int e0, int w,
                                        int r, int Z, F al, F be, F cap, F* ga, F* gb) {
  F q[MAXW], gq[MAXW];
  int vi[MAXW];
  bool qin[MAXW];   // QT: the V2C value was inside the int8 range (its clamp passes gradient)
  F m1 = F(INFINITY), m2 = F(INFINITY);
  int im = -1, im2 = -1, sgn = 0;
#pragma unroll
  for (int j = 0; j < MAXW; ++j) {
    if (j < w) {
      vi[j] = edge_var(edges[e0 + j], r, Z);
      const F xq = Qs[(e0 + j) * Z + r];
      q[j] = QT ? sat8f(xq) : xq;
      qin[j] = !QT || (xq >= F(-128) && xq <= F(127));
      const F A = mag_a<F, QT>(q[j], be);
      sgn ^= sbit(q[j]);
      if (A < m1) { m2 = m1; im2 = im; m1 = A; im = j; }
      else if (A < m2) { m2 = A; im2 = j; }
    }
  }
  // every edge other than the argmin took its magnitude from the argmin, the argmin from
  // the second minimum
  const bool cp1 = cap < m1, cp2 = cap < m2;
  const F c1 = cp1 ? cap : m1, c2 = cp2 ? cap : m2;
  F S = F(0), gmi = F(0);
#pragma unroll
  for (int j = 0; j < MAXW; ++j) {
    if (j < w) {
      const int ce = (e0 + j) * Z + r;
      const F s = (sgn ^ sbit(q[j])) ? F(-1) : F(1);
      F gv = gP[vi[j]];
      if constexpr (QT) {   // the posterior's saturation
        const F y = q[j] + s * (al * ((j == im) ? c2 : c1));
        if (y < F(-128) || y > F(127)) gv = F(0);
      }
      const F gm = s * (gv + gC[ce]);
      gq[j] = gv;
      if (j == im) gmi = gm;
      else S += gm;
    }
  }
  *ga = c1 * S + c2 * gmi;
  const F g_im = cp1 ? F(0) : al * S, g_im2 = cp2 ? F(0) : al * gmi;
  // |q| - beta > 0 there (without an argmin every magnitude is infinite, and capped)
  const bool act1 = m1 > F(0), act2 = m2 > F(0);
#pragma unroll
  for (int j = 0; j < MAXW; ++j) {
    if (j < w) {
      F g = gq[j];
      if (j == im && act1) g += sbit(q[j]) ? -g_im : g_im;
      if (j == im2 && act2) g += sbit(q[j]) ? -g_im2 : g_im2;
      g = qin[j] ? g : F(0);
      gP[vi[j]] = g;
      gC[(e0 + j) * Z + r] = -g;
    }
  }
  *gb = -((act1 ? g_im : F(0)) + (act2 ? g_im2 : F(0)));
}




template <typename F, bool GP, int MAXW, bool QT>
__global__ void __launch_bounds__(256) k_train_fwd(const TrainArgs a) {
  HIP_DYNAMIC_SHARED(char, smem);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t* tab = reinterpret_cast<uint32_t*>(smem);
  tables_to_lds(tab, a.edges, a.nedge, a.layer_w, a.mrow);
  __syncthreads();   // before any wave leaves
  const uint32_t* edges = tab;
  const int32_t* layer_w = reinterpret_cast<const int32_t*>(tab + a.nedge);
  const uint64_t cw = (uint64_t)blockIdx.x * a.waves + wave;
  if (cw >= a.batch) return;   // whole wave
  const TrainView<F> tv = train_view<F, GP>(a, cw, smem, wave);
  F* P = tv.P;
  const int N = a.n, E = a.e, Z = a.z, T = a.iters;
  const F* llr = reinterpret_cast<const F*>(a.llr) + cw * (uint64_t)N;
  const int K = a.seg, NS = (T + K - 1) / K;
  F* ckP = reinterpret_cast<F*>(a.ckP) + cw * (uint64_t)NS * N;
  F* ckC = reinterpret_cast<F*>(a.ckC) + cw * (uint64_t)NS * E;
  F* C = reinterpret_cast<F*>(a.wc) + cw * (uint64_t)E;   // working C2V, updated in place
  const F* alpha = reinterpret_cast<const F*>(a.alpha);
  const F* beta = reinterpret_cast<const F*>(a.beta);
  const F cap = (F)a.cap;
  for (int i = lane; i < N; i += 64) P[i] = llr[i];
  for (int i = lane; i < E; i += 64) C[i] = F(0);
  wave_sync();
  const int tlast = seg_start(NS - 1, T, K, NS);   // the adjoint starts with this segment: its q are kept now
  const int f = seg_start(1, T, K, NS);           // first segment's length (T when NS = 1)
  F* Qseg = reinterpret_cast<F*>(a.qs) + cw * (uint64_t)K * E;
  for (int t = 0; t < T; ++t) {
    F* Qs = t >= tlast ? Qseg + (uint64_t)(t - tlast) * E : (F*)nullptr;
    const int s = t == 0 ? 0 : (t >= f && (t - f) % K == 0 ? (t - f) / K + 1 : -1);
    if (s >= 0) {   // checkpoint at the start of segment s
      F* cp = ckP + (uint64_t)s * N;
      F* cc = ckC + (uint64_t)s * E;
      for (int i = lane; i < N; i += 64) cp[i] = P[i];
      for (int i = lane; i < E; i += 64) cc[i] = C[i];
    }
    int e0 = 0;
    for (int l = 0; l < a.mrow; ++l) {
      const int w = layer_w[l];
      const F al = alpha[t * a.mrow + l], be = offset_of<F, QT>(beta[t * a.mrow + l]);
      for (int r = lane; r < Z; r += 64) fwd_row<F, MAXW, QT>(a, edges, P, C, C, Qs, e0, w, r, al, be, cap);
      e0 += w;
      wave_sync();
    }
  }
  F* out = reinterpret_cast<F*>(a.post) + cw * (uint64_t)N;
  for (int i = lane; i < N; i += 64) out[i] = P[i];
}

template <typename F, bool GP, int MAXW, bool QT>
__global__ void __launch_bounds__(256) k_train_bwd(const TrainArgs a) {
  HIP_DYNAMIC_SHARED(char, smem);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t* tab = reinterpret_cast<uint32_t*>(smem);
  tables_to_lds(tab, a.edges, a.nedge, a.layer_w, a.mrow);
  __syncthreads();   // before any wave leaves
  const uint32_t* edges = tab;
  const int32_t* layer_w = reinterpret_cast<const int32_t*>(tab + a.nedge);
  const uint64_t cw = (uint64_t)blockIdx.x * a.waves + wave;
  if (cw >= a.batch) return;   // whole wave
  const TrainView<F> tv = train_view<F, GP>(a, cw, smem, wave);
  F* P = tv.P;
  F* gP = tv.gP;
  const int N = a.n, E = a.e, Z = a.z, T = a.iters, Lr = a.mrow;
  const int K = a.seg, NS = (T + K - 1) / K;
  const F* ckP = reinterpret_cast<const F*>(a.ckP) + cw * (uint64_t)NS * N;
  const F* ckC = reinterpret_cast<const F*>(a.ckC) + cw * (uint64_t)NS * E;
  F* C = reinterpret_cast<F*>(a.wc) + cw * (uint64_t)E;
  F* Qseg = reinterpret_cast<F*>(a.qs) + cw * (uint64_t)K * E;   // q of the segment's iterations
  F* gC = reinterpret_cast<F*>(a.gc) + cw * (uint64_t)E;
  F* part = reinterpret_cast<F*>(a.partial) + cw * (uint64_t)T * Lr * 2;
  const F* alpha = reinterpret_cast<const F*>(a.alpha);
  const F* beta = reinterpret_cast<const F*>(a.beta);
  const F cap = (F)a.cap;
  const F* gpost = reinterpret_cast<const F*>(a.grad_post) + cw * (uint64_t)N;
  for (int i = lane; i < N; i += 64) gP[i] = gpost[i];
  for (int i = lane; i < E; i += 64) gC[i] = F(0);
  wave_sync();
  for (int sg = NS - 1; sg >= 0; --sg) {
    const int t0 = seg_start(sg, T, K, NS), t1 = sg == NS - 1 ? T : seg_start(sg + 1, T, K, NS);
    // the segment's forward again, from its checkpoint, keeping every iteration's q (the
    // last segment's q were kept by the forward)
    if (sg < NS - 1) {
      const F* cp = ckP + (uint64_t)sg * N;
      const F* cc = ckC + (uint64_t)sg * E;
      for (int i = lane; i < N; i += 64) P[i] = cp[i];
      for (int i = lane; i < E; i += 64) C[i] = cc[i];
      wave_sync();
      for (int t = t0; t < t1; ++t) {
        F* Qs = Qseg + (uint64_t)(t - t0) * E;
        int e0 = 0;
        for (int l = 0; l < Lr; ++l) {
          const int w = layer_w[l];
          const F al = alpha[t * Lr + l], be = offset_of<F, QT>(beta[t * Lr + l]);
          for (int r = lane; r < Z; r += 64) fwd_row<F, MAXW, QT>(a, edges, P, C, C, Qs, e0, w, r, al, be, cap);
          e0 += w;
          wave_sync();
        }
      }
    }
    for (int t = t1 - 1; t >= t0; --t) {
      const F* Qs = Qseg + (uint64_t)(t - t0) * E;
      int e0 = a.nedge;
      for (int l = Lr - 1; l >= 0; --l) {
        const int w = layer_w[l];
        e0 -= w;
        const F al = alpha[t * Lr + l], be = offset_of<F, QT>(beta[t * Lr + l]);
        // per lane: the sum of its rows' sums, each row's terms summed from 0 in edge order
        F ga = F(0), gb = F(0);
        for (int r = lane; r < Z; r += 64) {
          F pa = F(0), pb = F(0);
          rev_row<F, MAXW, QT>(edges, gP, gC, Qs, e0, w, r, Z, al, be, cap, &pa, &pb);
          ga = r == lane ? pa : ga + pa;
          gb = r == lane ? pb : gb + pb;
        }
        wave_sync();
        ga = wave_sum(ga);
        gb = wave_sum(gb);
        if (lane == 0) {
          part[(t * Lr + l) * 2 + 0] = ga;
          part[(t * Lr + l) * 2 + 1] = gb;
        }
      }
    }
  }
  if (a.grad_llr) {
    F* gl = reinterpret_cast<F*>(a.grad_llr) + cw * (uint64_t)N;
    for (int i = lane; i < N; i += 64) gl[i] = gP[i];
  }
}

// out[k] = sum over codewords of partial[cw][k], k < T*mrow*2; block k sums one entry
template <typename F>
__global__ void __launch_bounds__(256) k_train_sum(const F* partial, uint32_t batch, int per_cw, F* galpha, F* gbeta) {
  __shared__ F red[4];
  const int k = blockIdx.x;
  F s = F(0);
  for (uint32_t c = threadIdx.x; c < batch; c += blockDim.x) s += partial[(uint64_t)c * per_cw + k];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    F t = F(0);
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
    if (k & 1) gbeta[k >> 1] = t;
    else galpha[k >> 1] = t;
  }
}

template <typename F, bool GP, int MAXW>
hipError_t launch_tw(const TrainArgs& a, int bwd, hipStream_t st) {
  const uint64_t blocks = (a.batch + a.waves - 1) / a.waves;
  const size_t lds = (size_t)a.tab_bytes + (GP ? 0 : (size_t)a.waves * 2 * a.n_pad * sizeof(F));
  if constexpr (std::is_same<F, float>::value) {
    if (a.quant) {
      if (bwd) hipLaunchKernelGGL((k_train_bwd<F, GP, MAXW, true>), dim3((uint32_t)blocks), dim3(64 * a.waves), lds, st, a);
      else hipLaunchKernelGGL((k_train_fwd<F, GP, MAXW, true>), dim3((uint32_t)blocks), dim3(64 * a.waves), lds, st, a);
      return hipGetLastError();
    }
  }
  if (bwd) hipLaunchKernelGGL((k_train_bwd<F, GP, MAXW, false>), dim3((uint32_t)blocks), dim3(64 * a.waves), lds, st, a);
  else hipLaunchKernelGGL((k_train_fwd<F, GP, MAXW, false>), dim3((uint32_t)blocks), dim3(64 * a.waves), lds, st, a);
  return hipGetLastError();
}

// the kernels' row loops unroll to the smallest weight bound that holds the code's rows
// (the register-C2V kernels of train_reg.hip when the code's shape is compiled for them)
template <typename F, bool GP>
hipError_t launch_t(const TrainArgs& a, int bwd, hipStream_t st) {
  hipError_t e = std::is_same<F, float>::value && a.reg_shape >= 0 ? launch_train_reg(a.reg_shape, a, bwd, st)
                 : a.maxw <= 8                                      ? launch_tw<F, GP, 8>(a, bwd, st)
                 : a.maxw <= 16                                     ? launch_tw<F, GP, 16>(a, bwd, st)
                 : a.maxw <= 24                                     ? launch_tw<F, GP, 24>(a, bwd, st)
                                                                    : launch_tw<F, GP, kTrainMaxW>(a, bwd, st);
  if (e != hipSuccess || !bwd) return e;
  const int per_cw = a.iters * a.mrow * 2;
  hipLaunchKernelGGL(k_train_sum<F>, dim3(per_cw), dim3(256), 0, st, reinterpret_cast<const F*>(a.partial),
                     (uint32_t)a.batch, per_cw, reinterpret_cast<F*>(a.grad_alpha), reinterpret_cast<F*>(a.grad_beta));
  return hipGetLastError();
}

template <typename F, bool GP, int MAXW>
const void* kernel_tw(const TrainArgs& a, int bwd) {
  if constexpr (std::is_same<F, float>::value) {
    if (a.quant)
      return bwd ? (const void*)k_train_bwd<F, GP, MAXW, true> : (const void*)k_train_fwd<F, GP, MAXW, true>;
  }
  return bwd ? (const void*)k_train_bwd<F, GP, MAXW, false> : (const void*)k_train_fwd<F, GP, MAXW, false>;
}

template <typename F, bool GP>
const void* kernel_t(const TrainArgs& a, int bwd) {
  return a.maxw <= 8    ? kernel_tw<F, GP, 8>(a, bwd)
         : a.maxw <= 16 ? kernel_tw<F, GP, 16>(a, bwd)
         : a.maxw <= 24 ? kernel_tw<F, GP, 24>(a, bwd)
                        : kernel_tw<F, GP, kTrainMaxW>(a, bwd);
}

}  // namespace

int train_resident_waves(const TrainArgs& a, int bwd) {
  const void* k = a.is_double ? (a.global_p ? kernel_t<double, true>(a, bwd) : kernel_t<double, false>(a, bwd))
                              : (a.global_p ? kernel_t<float, true>(a, bwd) : kernel_t<float, false>(a, bwd));
  const size_t lds = (size_t)a.tab_bytes +
                     (a.global_p ? 0 : (size_t)a.waves * 2 * a.n_pad * (a.is_double ? 8 : 4));
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64 * a.waves, lds) != hipSuccess) return 0;
  return per_cu * a.waves;
}

hipError_t launch_train(const TrainArgs& a, int bwd, hipStream_t st) {
  if (a.batch == 0) return hipSuccess;
  if (a.is_double) return a.global_p ? launch_t<double, true>(a, bwd, st) : launch_t<double, false>(a, bwd, st);
  return a.global_p ? launch_t<float, true>(a, bwd, st) : launch_t<float, false>(a, bwd, st);
}

}  // namespace fecsim
