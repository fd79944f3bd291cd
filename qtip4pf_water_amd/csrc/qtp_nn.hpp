// qtp_nn.hpp -- gfx950 device code of the learned centroid-force correction:
// an atom-centred network whose atomic energies sum to an additive
// correction V_corr of the qTIP4P/F potential, the place the reference's
// centroid hook takes (the splined pair corrections of forces.cpp:452-466,
// 634-691).  No counterpart in the reference otherwise.
//
// Descriptor of atom i (element e = 0 O, 1 H), positions in Angstrom:
//   G_m    = sum_j w[e_i][e_j] exp(-ETAR (r_ij - mu_m)^2) fc(r_ij; RR),
//            mu_m = MU0 + m DMU, m < 24                  (radial, r < 6 A)
//   G_24+n = 2^(1-z) sum_{j<k} w_ij w_ik (1 + l cos t_jik)^z
//            exp(-h (r_ij^2 + r_ik^2)) fc(r_ij; RA) fc(r_ik; RA)
//            (angular, r < 4 A; n = 8 hi + 2 zi + li: h = {0.02, 0.2},
//             z = {1, 2, 4, 8}, l = {+1, -1})
//   fc(r; R) = (cos(pi r / R) + 1) / 2
// (Behler-Parrinello symmetry functions with the element-pair weights of
// weighted ACSFs), normalised per element, g = (G - mean) / std.  Network per
// element: 40 -> 128 -> 128 -> 1, tanh, on BF16 matrix cores with FP32
// accumulation (v_mfma_f32_16x16x32_bf16).  Descriptors, forces and tangents
// are FP32; positions and the forces handed back are FP64.
#pragma once
#include "qtp_common.hpp"

namespace qtp {
namespace nn {

constexpr int NR = 24, NANG = 16, NIN = 40, NH = 128, KIN = 64;
constexpr int NMAX = 192;    // neighbours within RR per atom
constexpr int NAMAX = 64;    // of which within RA
constexpr int NPE = NH * NIN + NH + NH * NH + NH + NH + 1;  // parameters per element
constexpr float ANGF = 0.529177210903f;                       // Angstrom per bohr
constexpr double ANGD = 0.529177210903;
constexpr float RR = 6.0f, RA = 4.0f, MU0 = 0.8f, DMU = 5.2f / 23.0f;
constexpr float ETAR = 0.5f / (DMU * DMU);
constexpr float ETA_A0 = 0.02f, ETA_A1 = 0.2f;
constexpr float PIF = 3.14159265358979f;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float welem(int ei, int ej)
{
    // w[e_i][e_j]: O neighbours 1, H neighbours 1/2 (either centre)
    (void)ei;
    return ej == 0 ? 1.0f : 0.5f;
}
__device__ __forceinline__ int elem_of(int a) { return (a % 3) == 0 ? 0 : 1; }

// The value of lane (lane ^ o) of the wavefront (one ds_bpermute).
__device__ __forceinline__ float xor_lane(float v, int o)
{
    const int src = ((int)(threadIdx.x & 63) ^ o) << 2;
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(v)));
}

__device__ __forceinline__ float wave_sumf(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += xor_lane(v, o);
    return v;
}

// Frame f, atom a of the element-sorted row r (O rows first per element:
// O rows (f, mol), H rows (f, mol, h)).
__device__ __forceinline__ int row_atom(int e, int r, int n_mols, int n_frames, int n_atoms)
{
    if (e == 0) {
        if (r >= n_frames * n_mols) return -1;
        const int f = r / n_mols, m = r - f * n_mols;
        return f * n_atoms + 3 * m;
    }
    if (r >= 2 * n_frames * n_mols) return -1;
    const int f = r / (2 * n_mols), q = r - f * 2 * n_mols;
    return f * n_atoms + 3 * (q >> 1) + 1 + (q & 1);
}

struct NNGeom {
    int n_atoms, n_mols, n_frames, NAP;
    Box box;               // bohr
    float Lf[3], iLf[3];   // box edges in Angstrom (float) and their inverses
};

// Coordinate d of atom a in Angstrom (float), wrapped into the box first in
// FP64 (positions are never wrapped in the state).
__device__ __forceinline__ float posf(const double *__restrict__ X, int NAP, const Box &box, int d,
                                      int a)
{
    const double x = X[d * NAP + a];
    return (float)((x - box.L[d] * floor(x * box.inv[d])) * ANGD);
}
__device__ __forceinline__ float mimf(float d, float L, float iL) { return d - L * rintf(d * iL); }


// tanh(x) = 1 - 2 / (exp(2x) + 1): saturates to +-1 without overflow; its
// absolute error (a few 1e-7) is far below the BF16 rounding of the
// activations that follows.
__device__ __forceinline__ float tanh_fast(float x)
{
    return 1.0f - __fdividef(2.0f, __expf(2.0f * x) + 1.0f);
}

__device__ __forceinline__ float fcut(float r, float R)
{
    return r < R ? 0.5f * (__cosf(PIF * r / R) + 1.0f) : 0.0f;
}
__device__ __forceinline__ float dfcut(float r, float R)
{
    return r < R ? -0.5f * PIF / R * __sinf(PIF * r / R) : 0.0f;
}

// Per angular neighbour (r < RA): fc(r; RA), its derivative, and the two
// radial factors exp(-h r^2) of the angular terms (the pair factor
// exp(-h (r_ij^2 + r_ik^2)) is their product).
__device__ __forceinline__ float4 ang_pre(float r)
{
    return make_float4(fcut(r, RA), dfcut(r, RA), __expf(-ETA_A0 * r * r), __expf(-ETA_A1 * r * r));
}

// The 16 angular terms of one (j, k) pair (cos t = c, lengths ra, rb, their
// ang_pre values pa, pb, pair weight wt) and, when D, their partial
// derivatives in cos t, r_ij (Ta) and r_ik (Tb).
template <bool D>
__device__ __forceinline__ void ang_pair(float c, float ra, float rb, float4 pa, float4 pb, float wt,
                                         float *T, float *Tc, float *Ta, float *Tb)
{
    const float ff = pa.x * pb.x * wt;
    const float ex[2] = {pa.z * pb.z, pa.w * pb.w};
    const float eta[2] = {ETA_A0, ETA_A1};
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        const float base = l == 0 ? 1.0f + c : 1.0f - c;
        const float sg = l == 0 ? 1.0f : -1.0f;
        const float p1 = base, p2 = p1 * p1, p4 = p2 * p2, p8 = p4 * p4;
        // 2^(1-z) base^z and 2^(1-z) z base^(z-1)
        const float pw[4] = {p1, 0.5f * p2, 0.125f * p4, 0.0078125f * p8};
        const float dp[4] = {1.0f, p1, 0.5f * p2 * p1, 0.0625f * p4 * p2 * p1};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int z = 0; z < 4; ++z) {
                const int n = 8 * h + 2 * z + l;
                T[n] = pw[z] * ex[h] * ff;
                if (D) {
                    const float e = pw[z] * ex[h] * wt;
                    Tc[n] = sg * dp[z] * ex[h] * ff;
                    Ta[n] = e * pb.x * (pa.y - 2.0f * eta[h] * ra * pa.x);
                    Tb[n] = e * pa.x * (pb.y - 2.0f * eta[h] * rb * pb.x);
                }
            }
        }
    }
}

// Pair (j, k), j < k, of the upper triangle enumerated as p = k (k-1)/2 + j.
__device__ __forceinline__ void tri_decode(int p, int &j, int &k)
{
    int kk = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)p)) * 0.5f);
    while (kk * (kk - 1) / 2 > p) --kk;
    while ((kk + 1) * kk / 2 <= p) ++kk;
    k = kk;
    j = p - kk * (kk - 1) / 2;
}

// ---------------------------------------------------------------------------
// Per-atom kernels: one block of AT threads (four wavefronts) per atom, which
// share its neighbours, pairs and sums.  Positions come from k_nn_pos as
// float4 (Angstrom, wrapped into the box); separations are formed the same
// way everywhere (sep4), so lists, descriptors, forces and tangents see one
// geometry.
// ---------------------------------------------------------------------------
constexpr int AT = 256;
constexpr int RGRP = AT / NR;   // neighbour groups of the radial sums (10 x 24 threads)
constexpr int SEARCH_BATCH = 8; // candidate loads in flight per lane in k_nn_desc

// P4[f * n_atoms + a] = (x, y, z) of atom a of frame f in Angstrom, wrapped.
// grid: ceil(n_frames * n_atoms / 256), 256 threads.
__global__ __launch_bounds__(256) void k_nn_pos(NNGeom g, const double *__restrict__ pos,
                                                float4 *__restrict__ P4)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= g.n_frames * g.n_atoms) return;
    const int f = t / g.n_atoms, a = t - f * g.n_atoms;
    const double *X = pos + (long)f * 3 * g.NAP;
    P4[t] = make_float4(posf(X, g.NAP, g.box, 0, a), posf(X, g.NAP, g.box, 1, a),
                        posf(X, g.NAP, g.box, 2, a), 0.f);
}

__device__ __forceinline__ float4 sep4(const float4 *__restrict__ P, const NNGeom &g, int i, int j)
{
    const float4 a = P[i], b = P[j];
    const float dx = mimf(b.x - a.x, g.Lf[0], g.iLf[0]);
    const float dy = mimf(b.y - a.y, g.Lf[1], g.iLf[1]);
    const float dz = mimf(b.z - a.z, g.Lf[2], g.iLf[2]);
    return make_float4(dx, dy, dz, sqrtf(dx * dx + dy * dy + dz * dz));
}

// Block sum of n per-thread values (n <= NIN) into out[] (thread 0's view:
// every thread of the block may read out after the call).
template <int N>
__device__ __forceinline__ void block_sums(float *v, float (*red)[NIN], float *out)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        const float x = wave_sumf(v[m]);
        if (lane == m) red[w][m] = x;
    }
    __syncthreads();
    if (threadIdx.x < N) out[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] +
                                            red[2][threadIdx.x] + red[3][threadIdx.x];
    __syncthreads();
}

// Wave sums of 16 per-lane values by recursive halving: each round a lane
// keeps half of its values and adds its partner's matching half (8, 4, 2, 1
// values over the lane bits 5..2), then lanes 4c..4c+3 hold the total of
// component c = lane >> 2 after two plain xor rounds.  63 shuffles instead
// of 96, in a fixed order.
template <int H>
__device__ __forceinline__ void halving_round(float (&v)[16], int lane)
{
    constexpr int o = 4 * H;
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const float lo = v[i], hi = v[i + H];
        const float keep = up ? hi : lo, give = up ? lo : hi;
        v[i] = keep + xor_lane(give, o);
    }
}
__device__ __forceinline__ float halving_sum16(float (&v)[16])
{
    const int lane = threadIdx.x & 63;
    halving_round<8>(v, lane);
    halving_round<4>(v, lane);
    halving_round<2>(v, lane);
    halving_round<1>(v, lane);
    float x = v[0];
    x += xor_lane(x, 2);
    x += xor_lane(x, 1);
    return x;
}

// Neighbour lists and normalised descriptors.  The four wavefronts test
// consecutive quarters of the frame's atoms; their hits, merged in order,
// give lists in ascending index order: every atom within RR (nbr, count nn)
// and the subset within RA (nba, count nna).  status bit 2: a list was full
// (the atom's lists are then truncated).
// grid: n_frames * n_atoms blocks of AT threads.
__global__ __launch_bounds__(AT) void k_nn_desc(
    NNGeom g, const float4 *__restrict__ P4, const float *__restrict__ mean,
    const float *__restrict__ stdv, float *__restrict__ gt, int *__restrict__ nbr,
    int *__restrict__ nn, int *__restrict__ nba, int *__restrict__ nna, float *__restrict__ graw,
    int *__restrict__ status)
{
    __shared__ float4 tv[4][NMAX];
    __shared__ int tj[4][NMAX];
    __shared__ int wc[4];
    __shared__ float4 sv[NMAX];
    __shared__ int sj[NMAX];
    __shared__ float4 av[NAMAX], ap[NAMAX];
    __shared__ float aw[NAMAX];
    __shared__ int s_cnt, s_cnta;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int gi = blockIdx.x, f = gi / g.n_atoms, i = gi - f * g.n_atoms;
    const float4 *P = P4 + (long)f * g.n_atoms;
    const int ei = elem_of(i);
    const float rr2 = RR * RR;
    {
        const float4 pi = P[i];
        const int q = ((g.n_atoms + 255) / 256) * 64;
        const int lo = w * q, hi = min(g.n_atoms, lo + q);
        int c = 0;
        // SEARCH_BATCH candidates per lane loaded before any is tested, so
        // their latencies overlap
        for (int jb = lo; jb < hi; jb += 64 * SEARCH_BATCH) {
            float4 pb[SEARCH_BATCH];
#pragma unroll
            for (int u = 0; u < SEARCH_BATCH; ++u) {
                const int j = jb + 64 * u + lane;
                pb[u] = j < hi ? P[j] : pi;
            }
#pragma unroll
            for (int u = 0; u < SEARCH_BATCH; ++u) {
                if (jb + 64 * u >= hi) break;   // wave-uniform
                const int j = jb + 64 * u + lane;
                float dx = 0.f, dy = 0.f, dz = 0.f, r2 = 1e30f;
                if (j < hi && j != i) {
                    dx = mimf(pb[u].x - pi.x, g.Lf[0], g.iLf[0]);
                    dy = mimf(pb[u].y - pi.y, g.Lf[1], g.iLf[1]);
                    dz = mimf(pb[u].z - pi.z, g.Lf[2], g.iLf[2]);
                    r2 = dx * dx + dy * dy + dz * dz;
                }
                const bool in = r2 < rr2;
                const unsigned long long b = __ballot(in);
                const int s = c + __popcll(b & ((1ull << lane) - 1ull));
                if (in && s < NMAX) {
                    tv[w][s] = make_float4(dx, dy, dz, sqrtf(r2));
                    tj[w][s] = j;
                }
                c += __popcll(b);
            }
        }
        if (lane == 0) wc[w] = c;
    }
    __syncthreads();
    const int total = wc[0] + wc[1] + wc[2] + wc[3];
    const int cnt = min(total, NMAX);
    {
        int off = 0;
        for (int v = 0; v < w; ++v) off += wc[v];
        for (int s = lane; s < min(wc[w], NMAX); s += 64) {
            const int d = off + s;
            if (d < NMAX) {
                sv[d] = tv[w][s];
                sj[d] = tj[w][s];
                nbr[(long)gi * NMAX + d] = tj[w][s];
            }
        }
    }
    __syncthreads();
    if (w == 0) {
        int ca = 0;
        for (int s0 = 0; s0 < cnt; s0 += 64) {
            const int s = s0 + lane;
            const bool ina = s < cnt && sv[s].w < RA;
            const unsigned long long b = __ballot(ina);
            const int slot = ca + __popcll(b & ((1ull << lane) - 1ull));
            if (ina && slot < NAMAX) {
                const int j = sj[s];
                av[slot] = sv[s];
                ap[slot] = ang_pre(sv[s].w);
                aw[slot] = welem(ei, elem_of(j));
                nba[(long)gi * NAMAX + slot] = j;
            }
            ca += __popcll(b);
        }
        if (lane == 0) {
            s_cnt = cnt;
            s_cnta = min(ca, NAMAX);
            nn[gi] = cnt;
            nna[gi] = min(ca, NAMAX);
            if (total > NMAX || ca > NAMAX) atomicOr(status, 4);
        }
    }
    __syncthreads();
    // radial: the weight times the cutoff function per neighbour, then
    // thread (m, grp) sums component m over every RGRP-th neighbour
    __shared__ float rw[NMAX];
    __shared__ float rsum[RGRP][NR], asum[4][NANG];
    for (int s = tid; s < s_cnt; s += AT) rw[s] = welem(ei, elem_of(sj[s])) * fcut(sv[s].w, RR);
    __syncthreads();
    if (tid < RGRP * NR) {
        const int m = tid % NR, grp = tid / NR;
        const float mu = MU0 + m * DMU;
        float a = 0.f;
        for (int s = grp; s < s_cnt; s += RGRP) {
            const float x = sv[s].w - mu;
            a += rw[s] * __expf(-ETAR * x * x);
        }
        rsum[grp][m] = a;
    }
    // angular: a pair per thread, the 16 sums reduced per wave by halving
    {
        float T[NANG], acc[NANG];
#pragma unroll
        for (int n = 0; n < NANG; ++n) acc[n] = 0.f;
        const int na = s_cnta, npair = na * (na - 1) / 2;
        for (int p = tid; p < npair; p += AT) {
            int j, k;
            tri_decode(p, j, k);
            const float4 A = av[j], B = av[k];
            const float c = (A.x * B.x + A.y * B.y + A.z * B.z) / (A.w * B.w);
            ang_pair<false>(c, A.w, B.w, ap[j], ap[k], aw[j] * aw[k], T, nullptr, nullptr, nullptr);
#pragma unroll
            for (int n = 0; n < NANG; ++n) acc[n] += T[n];
        }
        const float x = halving_sum16(acc);
        if ((lane & 3) == 0) asum[w][lane >> 2] = x;
    }
    __syncthreads();
    if (tid < NIN) {
        float G = 0.f;
        if (tid < NR) {
#pragma unroll
            for (int grp = 0; grp < RGRP; ++grp) G += rsum[grp][tid];
        } else {
            G = asum[0][tid - NR] + asum[1][tid - NR] + asum[2][tid - NR] + asum[3][tid - NR];
        }
        gt[(long)gi * NIN + tid] = (G - mean[ei * NIN + tid]) / stdv[ei * NIN + tid];
        if (graw) graw[(long)gi * NIN + tid] = G;
    }
}

// ---------------------------------------------------------------------------
// BF16 matrix-core pieces.  v_mfma_f32_16x16x32_bf16: lane l holds
//   A: row l%16, k = 8 (l/16) + 0..7;  B: column l%16, k = 8 (l/16) + 0..7;
//   C: column l%16, rows 4 (l/16) + 0..3.
// Weights of one element, as BF16 rows of 16-byte groups:
//   w1 [128][64] (W1, inputs padded to 64), w1t [64][128] (its transpose),
//   w2 [128][128], w2t [128][128].
// ---------------------------------------------------------------------------
struct NNW {
    const __bf16 *w1, *w1t, *w2, *w2t;   // [2][...] per element
    const float *b1, *b2, *w3, *b3;       // [2][128], [2][128], [2][128], [2]
    float e_scale;
};
constexpr int W1E = NH * KIN, W2E = NH * NH;   // bf16 elements per element matrix

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 lds8(const __bf16 *p)
{
    return *reinterpret_cast<const bf16x8 *>(p);
}

// Load an A fragment of 16 rows x 32 k (k-step s) from a float matrix
// src[row][ld] (rows of the tile: atom index or -1), zero beyond kmax.
__device__ __forceinline__ bf16x8 afrag_f32(const float *__restrict__ src, long row, int ld,
                                            int k0, int kmax)
{
    bf16x8 a;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int k = k0 + t;
        a[t] = (__bf16)((row >= 0 && k < kmax) ? src[row * ld + k] : 0.0f);
    }
    return a;
}

// C-layout values (4 rows x 8 column tiles per lane) -> a 16 x 128 BF16 tile
// in LDS (row-major), then A fragments of it for k-step s.
__device__ __forceinline__ void c_to_lds(const float (*v)[4], __bf16 *tile, int lane)
{
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) tile[(4 * (lane >> 4) + r) * NH + 16 * t + (lane & 15)] = (__bf16)v[t][r];
}

__device__ __forceinline__ bf16x8 gload8(const __bf16 *p)
{
    return *reinterpret_cast<const bf16x8 *>(p);
}

// Per-atom energies and dE/dG (G the raw descriptor) of one 16-row tile of
// one element per block; its four wavefronts split every layer's 128 (or
// 48) output columns, exchanging the activations through LDS.  Every B
// fragment (weights, L2-resident) is loaded at the start, so the loads
// overlap each other and the first layer.  Blocks [0, nto) take O tiles, the
// rest H tiles.  Etile (may be null): the tile's energy sum per block.
// grid: nto + nth blocks, 256 threads.
constexpr int MLP_WAVES = 4;
__global__ __launch_bounds__(64 * MLP_WAVES) void k_nn_mlp(
    NNGeom g, NNW W, int nto, const float *__restrict__ gt, const float *__restrict__ stdv,
    float *__restrict__ dEdG, float *__restrict__ Eatom, float *__restrict__ Etile)
{
    __shared__ __align__(16) __bf16 sA[16 * NH];   // a1, later dE/dz1
    __shared__ __align__(16) __bf16 sD[16 * NH];   // dE/dz2
    __shared__ float sE[MLP_WAVES][16];
    const int e = blockIdx.x < nto ? 0 : 1;
    const int tile = e == 0 ? blockIdx.x : blockIdx.x - nto;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 15, kq = 8 * (lane >> 4);
    const int row0 = tile * 16, rq = 4 * (lane >> 4);
    const int ra = row_atom(e, row0 + col, g.n_mols, g.n_frames, g.n_atoms);
    const __bf16 *w1 = W.w1 + (long)e * W1E, *w1t = W.w1t + (long)e * W1E;
    const __bf16 *w2 = W.w2 + (long)e * W2E, *w2t = W.w2t + (long)e * W2E;
    bf16x8 B1[2][2], B2[4][2], B3[4][2], B4[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            const int n = 32 * w + 16 * tt + col;
            if (s < 2) B1[s][tt] = gload8(w1 + n * KIN + 32 * s + kq);
            B2[s][tt] = gload8(w2 + n * NH + 32 * s + kq);
            B3[s][tt] = gload8(w2t + n * NH + 32 * s + kq);
        }
        B4[s] = gload8(w1t + (16 * (w < 3 ? w : 0) + col) * NH + 32 * s + kq);
    }
    // layer 1, columns 32 w .. 32 w + 31
    float a1[2][4];
    {
        f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bf16x8 a = afrag_f32(gt, ra, NIN, 32 * s + kq, NIN);
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) acc[tt] = mfma16(a, B1[s][tt], acc[tt]);
        }
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            const int n = 32 * w + 16 * tt + col;
            const float b = W.b1[e * NH + n];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                a1[tt][r] = tanh_fast(acc[tt][r] + b);
                sA[(rq + r) * NH + n] = (__bf16)a1[tt][r];
            }
        }
    }
    __syncthreads();
    // layer 2, energy, dE/dz2
    {
        f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const bf16x8 a = lds8(sA + col * NH + 32 * s + kq);
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) acc[tt] = mfma16(a, B2[s][tt], acc[tt]);
        }
        float er[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            const int n = 32 * w + 16 * tt + col;
            const float b = W.b2[e * NH + n], w3 = W.w3[e * NH + n];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float a2 = tanh_fast(acc[tt][r] + b);
                er[r] += w3 * a2;
                sD[(rq + r) * NH + n] = (__bf16)(W.e_scale * w3 * (1.0f - a2 * a2));
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = er[r];
            v += xor_lane(v, 1);
            v += xor_lane(v, 2);
            v += xor_lane(v, 4);
            v += xor_lane(v, 8);
            if (col == 0) sE[w][rq + r] = v;
        }
    }
    __syncthreads();
    // dE/da1 for columns 32 w ..; dE/dz1 = (1 - a1^2) dE/da1
    {
        f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const bf16x8 a = lds8(sD + col * NH + 32 * s + kq);
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) acc[tt] = mfma16(a, B3[s][tt], acc[tt]);
        }
        // every wavefront has read sA (layer 2) before the barrier above
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            const int n = 32 * w + 16 * tt + col;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                sA[(rq + r) * NH + n] = (__bf16)((1.0f - a1[tt][r] * a1[tt][r]) * acc[tt][r]);
        }
    }
    if (w == 0 && lane < 16) {
        const int rr = row_atom(e, row0 + lane, g.n_mols, g.n_frames, g.n_atoms);
        float v = rr >= 0 ? W.e_scale * (sE[0][lane] + sE[1][lane] + sE[2][lane] + sE[3][lane] + W.b3[e])
                          : 0.f;
        if (rr >= 0 && Eatom) Eatom[rr] = v;
        v += xor_lane(v, 1);
        v += xor_lane(v, 2);
        v += xor_lane(v, 4);
        v += xor_lane(v, 8);
        if (lane == 0 && Etile) Etile[blockIdx.x] = v;
    }
    __syncthreads();
    // dE/dg = dE/dz1 W1: column tile w (w < 3: inputs 0..47), dE/dG = dE/dg / std
    if (w < 3) {
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mfma16(lds8(sA + col * NH + 32 * s + kq), B4[s], acc);
        const int d = 16 * w + col;
        if (d < NIN) {
            const float is = 1.0f / stdv[e * NIN + d];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = row_atom(e, row0 + rq + r, g.n_mols, g.n_frames, g.n_atoms);
                if (rr >= 0) dEdG[(long)rr * NIN + d] = acc[r] * is;
            }
        }
    }
}

// d(cos t)/dA and d(cos t)/dB for A = x_j - x_i, B = x_k - x_i.
__device__ __forceinline__ void dcos(float4 A, float4 B, float c, float3 &dA, float3 &dB)
{
    const float iab = 1.0f / (A.w * B.w), ia2 = c / (A.w * A.w), ib2 = c / (B.w * B.w);
    dA = make_float3(B.x * iab - ia2 * A.x, B.y * iab - ia2 * A.y, B.z * iab - ia2 * A.z);
    dB = make_float3(A.x * iab - ib2 * B.x, A.y * iab - ib2 * B.y, A.z * iab - ib2 * B.z);
}


// Radial derivative d phi_m / dr of exp(-ETAR (r - mu_m)^2) fc(r; RR).
__device__ __forceinline__ float dphi(int m, float r, float fc, float dfc)
{
    const float x = r - (MU0 + m * DMU);
    return __expf(-ETAR * x * x) * (dfc - 2.0f * ETAR * x * fc);
}

// ---------------------------------------------------------------------------
// Angular force pieces per centre i: for every angular neighbour j (slot of
// nba), D_i[j] = sum_{k != j} dE_i/dx_j of the pair terms T(i; j, k), and
// Dself_i = dE_i/dx_i|angular = -sum_j D_i[j].  Each unordered pair (j, k)
// is differentiated once, by one thread, for both members; the pieces are
// summed in LDS as 64-bit fixed-point integers (resolution 2^-40, range
// +-2^23), whose additions are exact, so the sums do not depend on the order
// in which the threads add them (nor on which member is called j).
// grid: n_frames * n_atoms blocks of AT threads.
// ---------------------------------------------------------------------------
constexpr double DANG_FIX = 1099511627776.0;   // 2^40
__device__ __forceinline__ void fix_add(unsigned long long *a, float v)
{
    atomicAdd(a, (unsigned long long)(long long)llrint((double)v * DANG_FIX));
}
__global__ __launch_bounds__(AT) void k_nn_dang(
    NNGeom g, const float4 *__restrict__ P4, const int *__restrict__ nba,
    const int *__restrict__ nna, const float *__restrict__ dEdG, float4 *__restrict__ Dang,
    float4 *__restrict__ Dself)
{
    __shared__ float4 av[NAMAX], ap[NAMAX];
    __shared__ float aw[NAMAX], se[NANG];
    __shared__ unsigned long long acc[3][NAMAX];
    const int tid = threadIdx.x;
    const int gi = blockIdx.x, f = gi / g.n_atoms, i = gi - f * g.n_atoms;
    const float4 *P = P4 + (long)f * g.n_atoms;
    const int ei = elem_of(i), cnta = nna[gi];
    if (tid < cnta) {
        const int j = nba[(long)gi * NAMAX + tid];
        const float4 A = sep4(P, g, i, j);
        av[tid] = A;
        ap[tid] = ang_pre(A.w);
        aw[tid] = welem(ei, elem_of(j));
    }
    if (tid < NANG) se[tid] = dEdG[(long)gi * NIN + NR + tid];
    if (tid < 3 * NAMAX) (&acc[0][0])[tid] = 0ull;
    __syncthreads();
    // pairs in circulant order: p = n (d - 1) + j gives (j, (j + d) mod n),
    // d = 1 .. n/2 (for even n only j < n/2 at d = n/2): every unordered pair
    // once, and the lanes of a wavefront mostly hold distinct j and distinct
    // k, so their atomic adds do not pile onto one address
    const int npair = cnta * (cnta - 1) / 2;
    for (int p = tid; p < npair; p += AT) {
        const int d = p / cnta + 1, j = p - (d - 1) * cnta;
        const int k = j + d < cnta ? j + d : j + d - cnta;
        const float4 A = av[j], B = av[k];
        float T[NANG], Tc[NANG], Ta[NANG], Tb[NANG];
        const float iab = 1.0f / (A.w * B.w);
        const float c = (A.x * B.x + A.y * B.y + A.z * B.z) * iab;
        ang_pair<true>(c, A.w, B.w, ap[j], ap[k], aw[j] * aw[k], T, Tc, Ta, Tb);
        float Sc = 0.f, Sa = 0.f, Sb = 0.f;
#pragma unroll
        for (int n = 0; n < NANG; ++n) {
            Sc += se[n] * Tc[n];
            Sa += se[n] * Ta[n];
            Sb += se[n] * Tb[n];
        }
        // d cos / dA = B / (ra rb) - c A / ra^2, and likewise for B
        const float ca = Sc * iab;
        const float cj = Sa / A.w - Sc * c / (A.w * A.w);
        const float ck = Sb / B.w - Sc * c / (B.w * B.w);
        fix_add(&acc[0][j], ca * B.x + cj * A.x);
        fix_add(&acc[1][j], ca * B.y + cj * A.y);
        fix_add(&acc[2][j], ca * B.z + cj * A.z);
        fix_add(&acc[0][k], ca * A.x + ck * B.x);
        fix_add(&acc[1][k], ca * A.y + ck * B.y);
        fix_add(&acc[2][k], ca * A.z + ck * B.z);
    }
    __syncthreads();
    if (tid < 64) {
        float4 D = make_float4(0.f, 0.f, 0.f, 0.f);
        if (tid < cnta) {
            D.x = (float)((double)(long long)acc[0][tid] / DANG_FIX);
            D.y = (float)((double)(long long)acc[1][tid] / DANG_FIX);
            D.z = (float)((double)(long long)acc[2][tid] / DANG_FIX);
            Dang[(long)gi * NAMAX + tid] = D;
        }
        const float sx = wave_sumf(D.x), sy = wave_sumf(D.y), sz = wave_sumf(D.z);
        if (tid == 0) Dself[gi] = make_float4(-sx, -sy, -sz, 0.f);
    }
}

// ---------------------------------------------------------------------------
// Forces of the correction, F_a = -dE/dx_a, gathered per atom a (fixed
// summation order, no atomics):
//   radial:  dE/dx_a = -sum_j [sum_m (e_am w_aj + e_jm w_ja) phi'_m(r_aj)] u_aj
//   angular: Dself_a + sum over angular neighbours i of D_i[slot of a in i]
// with e = dE/dG and u_aj the unit vector from a to j.  MODE 0 writes F to an
// FP64 planar force array; MODE 1 writes the residual F - target (training)
// and the atom's squared residual (and F to fout when given).
// grid: n_frames * n_atoms blocks of AT threads.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(AT) void k_nn_force(
    NNGeom g, const float4 *__restrict__ P4, const int *__restrict__ nbr,
    const int *__restrict__ nn, const int *__restrict__ nba, const int *__restrict__ nna,
    const float *__restrict__ dEdG, const float4 *__restrict__ Dang,
    const float4 *__restrict__ Dself, double *__restrict__ fout,
    const double *__restrict__ target, float *__restrict__ res, float *__restrict__ sq)
{
    __shared__ float se[NR];
    __shared__ float red[4][NIN], tot[NIN];
    const int tid = threadIdx.x;
    const int ga = blockIdx.x, f = ga / g.n_atoms, a = ga - f * g.n_atoms, base = f * g.n_atoms;
    const float4 *P = P4 + base;
    const int ea = elem_of(a);
    if (tid < NR) se[tid] = dEdG[(long)ga * NIN + tid];
    __syncthreads();
    const int cnt = nn[ga], cnta = nna[ga];
    float v[3] = {0.f, 0.f, 0.f};   // dE/dx_a (Angstrom)
    // radial: thread pair (2s, 2s + 1) takes neighbour s, each half of the
    // 24 components
    for (int t = tid; t < 2 * cnt; t += AT) {
        const int s = t >> 1, m0 = (t & 1) * (NR / 2);
        const int j = nbr[(long)ga * NMAX + s];
        const float4 d = sep4(P, g, a, j);
        const float fc = fcut(d.w, RR), dfc = dfcut(d.w, RR);
        const int ej = elem_of(j);
        const float waj = welem(ea, ej), wja = welem(ej, ea);
        const float *ejv = dEdG + (long)(base + j) * NIN;
        float c = 0.f;
#pragma unroll
        for (int mm = 0; mm < NR / 2; ++mm) {
            const int m = m0 + mm;
            c += (se[m] * waj + ejv[m] * wja) * dphi(m, d.w, fc, dfc);
        }
        const float ir = c / d.w;
        v[0] -= ir * d.x;
        v[1] -= ir * d.y;
        v[2] -= ir * d.z;
    }
    if (tid < cnta) {
        // a in the (ascending) angular list of its neighbour i
        const int gi = base + nba[(long)ga * NAMAX + tid];
        const int *L = nba + (long)gi * NAMAX;
        const int ni = nna[gi];
        int lo = 0, hi = ni;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (L[mid] < a) lo = mid + 1;
            else hi = mid;
        }
        if (lo < ni && L[lo] == a) {
            const float4 D = Dang[(long)gi * NAMAX + lo];
            v[0] += D.x;
            v[1] += D.y;
            v[2] += D.z;
        }
    }
    if (tid == 0) {
        const float4 D = Dself[ga];
        v[0] += D.x;
        v[1] += D.y;
        v[2] += D.z;
    }
    block_sums<3>(v, red, tot);
    if (tid == 0) {
        const double F[3] = {-ANGD * tot[0], -ANGD * tot[1], -ANGD * tot[2]};
        double *fo = fout ? fout + (long)f * 3 * g.NAP : nullptr;
        if (MODE == 0) {
            for (int d = 0; d < 3; ++d) fo[d * g.NAP + a] = F[d];
        } else {
            const double *tg = target + (long)f * 3 * g.NAP;
            float s2 = 0.f;
            for (int d = 0; d < 3; ++d) {
                const float r = (float)(F[d] - tg[d * g.NAP + a]);
                res[(long)ga * 3 + d] = r;
                s2 += r * r;
                if (fo) fo[d * g.NAP + a] = F[d];
            }
            sq[ga] = s2;
        }
    }
}

// ---------------------------------------------------------------------------
// Tangent of the normalised descriptor along the residual field R (training):
//   u_i = ANG * (dG_i/dx . R) / std     (R per atom, Hartree/bohr)
// radial:  sum_j w phi'_m(r_ij) u_ij . (R_j - R_i)
// angular: sum_{j<k} Tc dc + Ta da + Tb db, dc = dcos/dA . dA + dcos/dB . dB,
//          da = A . (R_j - R_i) / r_ij, db = B . (R_k - R_i) / r_ik
// grid: n_frames * n_atoms blocks of AT threads.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(AT) void k_nn_tan(
    NNGeom g, const float4 *__restrict__ P4, const int *__restrict__ nbr,
    const int *__restrict__ nn, const int *__restrict__ nba, const int *__restrict__ nna,
    const float *__restrict__ res, const float *__restrict__ stdv, float *__restrict__ ut)
{
    __shared__ float4 av[NAMAX], ap[NAMAX], ar[NAMAX];   // vectors, ang_pre, R_j - R_i
    __shared__ float aw[NAMAX];
    __shared__ float red[4][NIN], su[NIN];
    const int tid = threadIdx.x;
    const int gi = blockIdx.x, f = gi / g.n_atoms, i = gi - f * g.n_atoms, base = f * g.n_atoms;
    const float4 *P = P4 + base;
    const int ei = elem_of(i);
    const float Rix = res[(long)gi * 3], Riy = res[(long)gi * 3 + 1], Riz = res[(long)gi * 3 + 2];
    const int cnt = nn[gi], cnta = nna[gi];
    if (tid < cnta) {
        const int j = nba[(long)gi * NAMAX + tid];
        const float4 A = sep4(P, g, i, j);
        av[tid] = A;
        ap[tid] = ang_pre(A.w);
        aw[tid] = welem(ei, elem_of(j));
        const long o = (long)(base + j) * 3;
        ar[tid] = make_float4(res[o] - Rix, res[o + 1] - Riy, res[o + 2] - Riz, 0.f);
    }
    __syncthreads();
    float acc[NIN];
#pragma unroll
    for (int m = 0; m < NIN; ++m) acc[m] = 0.f;
    for (int s = tid; s < cnt; s += AT) {
        const int j = nbr[(long)gi * NMAX + s];
        const float4 d = sep4(P, g, i, j);
        const long o = (long)(base + j) * 3;
        const float dr = (d.x * (res[o] - Rix) + d.y * (res[o + 1] - Riy) + d.z * (res[o + 2] - Riz)) / d.w;
        const float wf = welem(ei, elem_of(j)) * dr;
        const float fc = fcut(d.w, RR), dfc = dfcut(d.w, RR);
#pragma unroll
        for (int m = 0; m < NR; ++m) acc[m] += wf * dphi(m, d.w, fc, dfc);
    }
    {
        float T[NANG], Tc[NANG], Ta[NANG], Tb[NANG];
        const int npair = cnta * (cnta - 1) / 2;
        for (int p = tid; p < npair; p += AT) {
            int j, k;
            tri_decode(p, j, k);
            const float4 A = av[j], B = av[k], dA4 = ar[j], dB4 = ar[k];
            const float c = (A.x * B.x + A.y * B.y + A.z * B.z) / (A.w * B.w);
            ang_pair<true>(c, A.w, B.w, ap[j], ap[k], aw[j] * aw[k], T, Tc, Ta, Tb);
            float3 cA, cB;
            dcos(A, B, c, cA, cB);
            const float dc = cA.x * dA4.x + cA.y * dA4.y + cA.z * dA4.z + cB.x * dB4.x + cB.y * dB4.y +
                             cB.z * dB4.z;
            const float da = (A.x * dA4.x + A.y * dA4.y + A.z * dA4.z) / A.w;
            const float db = (B.x * dB4.x + B.y * dB4.y + B.z * dB4.z) / B.w;
#pragma unroll
            for (int n = 0; n < NANG; ++n) acc[NR + n] += Tc[n] * dc + Ta[n] * da + Tb[n] * db;
        }
    }
    block_sums<NIN>(acc, red, su);
    if (tid < NIN) ut[(long)gi * NIN + tid] = ANGF * su[tid] / stdv[ei * NIN + tid];
}

// Sum of per-atom values in a fixed order (one block): out[0] = scale * sum
// (+ out[0] when accumulate).  grid: 1, 1024 threads.
__global__ __launch_bounds__(1024) void k_nn_sum(int n, const float *__restrict__ v, double scale,
                                                 int accumulate, double *__restrict__ out)
{
    __shared__ double sh[16];
    double s = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) s += (double)v[t];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[0] = (accumulate ? out[0] : 0.0) + scale * s;
}

// ---------------------------------------------------------------------------
// Training, forward-over-reverse on 16-row tiles (one wavefront per tile):
// with z1 = W1 g + b1, a1 = tanh z1, s1 = 1 - a1^2, the tangent zd1 = W1 u,
// ad1 = s1 zd1, likewise z2, a2, s2, zd2, ad2 = s2 zd2, and
// J = e_scale sum_rows w3 . ad2, the adjoints
//   d_ad2 = e_scale w3, d_zd2 = s2 d_ad2, d_z2 = -2 a2 s2 zd2 d_ad2,
//   d_ad1 = d_zd2 W2, d_a1 = d_z2 W2 - 2 a1 zd1 d_ad1, d_zd1 = s1 d_ad1,
//   d_z1 = s1 d_a1
// give dJ/dW2 = sum d_zd2 (x) ad1 + d_z2 (x) a1, dJ/dW1 = sum d_zd1 (x) u +
// d_z1 (x) g, dJ/db = sum d_z, dJ/dw3 = e_scale sum ad2.  The weight
// gradients are left for k_nn_wgrad as transposed BF16 matrices [unit][row]
// per element (rows padded to WG_ROWS); bias and w3 parts as per-tile sums.
// grid: nbo + nbh blocks (each element's tiles padded to TRAIN_WAVES per
// block, O first), 64 * TRAIN_WAVES threads, TRAIN_LDS bytes of LDS.
// ---------------------------------------------------------------------------
constexpr int WG_ROWS = 256;
constexpr int TRAIN_WAVES = 4;   // 16-row tiles (one per wavefront) per training block
constexpr int TRAIN_LDS = 2 * (W1E + 2 * W2E) + TRAIN_WAVES * 2 * 16 * NH * 2;

struct NNTrainBuf {
    __bf16 *d2t[2], *d2p[2], *a1t[2], *a1p[2], *d1t[2], *d1p[2], *a0t[2], *a0p[2];
    int rows[2];          // padded row count per element (leading dimension)
    float *tp;            // [tile][3][128]: d_z1, d_z2 sums, e_scale ad2 sums
};

__device__ __forceinline__ void store_t(__bf16 *dst, int ld, int unit, int row, const float *v4)
{
    // four consecutive rows of one unit: one 8-byte store
    __bf16 b[4] = {(__bf16)v4[0], (__bf16)v4[1], (__bf16)v4[2], (__bf16)v4[3]};
    *reinterpret_cast<uint2 *>(dst + (long)unit * ld + row) = *reinterpret_cast<uint2 *>(b);
}

__global__ __launch_bounds__(64 * TRAIN_WAVES) void k_nn_train_fwd(
    NNGeom g, NNW W, int nbo, const float *__restrict__ gt, const float *__restrict__ ut,
    NNTrainBuf B)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __bf16 *s_w1 = reinterpret_cast<__bf16 *>(smem);
    __bf16 *s_w2 = s_w1 + W1E;
    __bf16 *s_w2t = s_w2 + W2E;
    __bf16 *s_tile = s_w2t + W2E;
    const int e = blockIdx.x < nbo ? 0 : 1;
    const int tile = (e == 0 ? blockIdx.x : blockIdx.x - nbo) * TRAIN_WAVES + (threadIdx.x >> 6);
    const int gtile = blockIdx.x * TRAIN_WAVES + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, col = lane & 15, kq = 8 * (lane >> 4);
    __bf16 *tA = s_tile + (threadIdx.x >> 6) * 2 * 16 * NH, *tB = tA + 16 * NH;
    {
        const uint4 *g1 = reinterpret_cast<const uint4 *>(W.w1 + (long)e * W1E);
        const uint4 *g2 = reinterpret_cast<const uint4 *>(W.w2 + (long)e * W2E);
        const uint4 *g2t = reinterpret_cast<const uint4 *>(W.w2t + (long)e * W2E);
        uint4 *d1 = reinterpret_cast<uint4 *>(s_w1), *d2 = reinterpret_cast<uint4 *>(s_w2);
        uint4 *d2t = reinterpret_cast<uint4 *>(s_w2t);
        for (int t = threadIdx.x; t < W1E / 8; t += blockDim.x) d1[t] = g1[t];
        for (int t = threadIdx.x; t < W2E / 8; t += blockDim.x) {
            d2[t] = g2[t];
            d2t[t] = g2t[t];
        }
    }
    __syncthreads();
    const int ld = B.rows[e];
    const int row0 = tile * 16;
    const int ra = row_atom(e, row0 + col, g.n_mols, g.n_frames, g.n_atoms);
    const int rq = row0 + 4 * (lane >> 4);   // first of this lane's four C rows
    // layer 1, primal and tangent; the inputs go to a0p / a0t
    float a1[8][4], zd1[8][4];
    {
        f32x4 az[8], ad[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) az[t] = ad[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bf16x8 ag = afrag_f32(gt, ra, NIN, 32 * s + kq, NIN);
            const bf16x8 au = afrag_f32(ut, ra, NIN, 32 * s + kq, NIN);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const bf16x8 b = lds8(s_w1 + (16 * t + col) * KIN + 32 * s + kq);
                az[t] = mfma16(ag, b, az[t]);
                ad[t] = mfma16(au, b, ad[t]);
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                B.a0p[e][(long)(32 * s + kq + q) * ld + row0 + col] = ag[q];
                B.a0t[e][(long)(32 * s + kq + q) * ld + row0 + col] = au[q];
            }
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const float b = W.b1[e * NH + 16 * t + col];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                a1[t][r] = tanh_fast(az[t][r] + b);
                zd1[t][r] = ad[t][r];
            }
        }
    }
    {
        float ad1[8][4];
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) ad1[t][r] = (1.0f - a1[t][r] * a1[t][r]) * zd1[t][r];
        c_to_lds(a1, tA, lane);
        c_to_lds(ad1, tB, lane);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            store_t(B.a1p[e], ld, 16 * t + col, rq, a1[t]);
            store_t(B.a1t[e], ld, 16 * t + col, rq, ad1[t]);
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // layer 2, primal and tangent; adjoints of the output
    float dzd2[8][4], dz2[8][4];
    {
        f32x4 az[8], ad[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) az[t] = ad[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const bf16x8 xa = lds8(tA + col * NH + 32 * s + kq);
            const bf16x8 xd = lds8(tB + col * NH + 32 * s + kq);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const bf16x8 b = lds8(s_w2 + (16 * t + col) * NH + 32 * s + kq);
                az[t] = mfma16(xa, b, az[t]);
                ad[t] = mfma16(xd, b, ad[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int n = 16 * t + col;
            const float b = W.b2[e * NH + n], dad2 = W.e_scale * W.w3[e * NH + n];
            float sz = 0.f, sw = 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float a2 = tanh_fast(az[t][r] + b), s2 = 1.0f - a2 * a2;
                const float zd2 = ad[t][r];
                dzd2[t][r] = s2 * dad2;
                dz2[t][r] = -2.0f * a2 * s2 * zd2 * dad2;
                sz += dz2[t][r];
                sw += W.e_scale * s2 * zd2;
            }
            sz += xor_lane(sz, 16);
            sz += xor_lane(sz, 32);
            sw += __shfl_xor(sw, 16, 64);
            sw += __shfl_xor(sw, 32, 64);
            if (lane < 16) {
                B.tp[((long)gtile * 3 + 1) * NH + n] = sz;
                B.tp[((long)gtile * 3 + 2) * NH + n] = sw;
            }
            store_t(B.d2t[e], ld, n, rq, dzd2[t]);
            store_t(B.d2p[e], ld, n, rq, dz2[t]);
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    c_to_lds(dzd2, tA, lane);
    c_to_lds(dz2, tB, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    {
        f32x4 gd[8], gp[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) gd[t] = gp[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const bf16x8 xd = lds8(tA + col * NH + 32 * s + kq);
            const bf16x8 xp = lds8(tB + col * NH + 32 * s + kq);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const bf16x8 b = lds8(s_w2t + (16 * t + col) * NH + 32 * s + kq);
                gd[t] = mfma16(xd, b, gd[t]);
                gp[t] = mfma16(xp, b, gp[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int n = 16 * t + col;
            float dzd1[4], dz1[4], sz = 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float s1 = 1.0f - a1[t][r] * a1[t][r];
                const float dad1 = gd[t][r];
                const float da1 = gp[t][r] - 2.0f * a1[t][r] * zd1[t][r] * dad1;
                dzd1[r] = s1 * dad1;
                dz1[r] = s1 * da1;
                sz += dz1[r];
            }
            sz += xor_lane(sz, 16);
            sz += xor_lane(sz, 32);
            if (lane < 16) B.tp[((long)gtile * 3 + 0) * NH + n] = sz;
            store_t(B.d1t[e], ld, n, rq, dzd1);
            store_t(B.d1p[e], ld, n, rq, dz1);
        }
    }
}

// Weight gradients as split-K products over WG_ROWS-row chunks of one
// element: part[e][chunk] = [dJ/dW1 (128 x 64) | dJ/dW2 (128 x 128)],
//   dJ/dW2[m][n] = sum_rows d2t[m] a1t[n] + d2p[m] a1p[n]
//   dJ/dW1[m][k] = sum_rows d1t[m] a0t[k] + d1p[m] a0p[k]
// 88 output tiles of 16 x 16 (64 of W2, 24 of W1 for k < 48) over 8 waves.
// grid: nco + nch blocks (chunks of O, then of H), 512 threads.
constexpr int WG_WAVES = 8;
constexpr int WG_PART = NH * KIN + NH * NH;
__global__ __launch_bounds__(64 * WG_WAVES) void k_nn_wgrad(NNTrainBuf B, int nco,
                                                            float *__restrict__ part)
{
    const int e = blockIdx.x < nco ? 0 : 1;
    const int chunk = e == 0 ? blockIdx.x : blockIdx.x - nco;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 15, kq = 8 * (lane >> 4);
    const int ld = B.rows[e];
    const long r0 = (long)chunk * WG_ROWS;
    float *P = part + (long)blockIdx.x * WG_PART;
    for (int tt = wv; tt < 88; tt += WG_WAVES) {
        const bool w2 = tt < 64;
        const int mt = w2 ? tt >> 3 : (tt - 64) / 3, nt = w2 ? tt & 7 : (tt - 64) % 3;
        const __bf16 *At = w2 ? B.d2t[e] : B.d1t[e], *Ap = w2 ? B.d2p[e] : B.d1p[e];
        const __bf16 *Bt = w2 ? B.a1t[e] : B.a0t[e], *Bp = w2 ? B.a1p[e] : B.a0p[e];
        const long ra = (long)(16 * mt + col) * ld + r0 + kq, rb = (long)(16 * nt + col) * ld + r0 + kq;
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
        for (int s = 0; s < WG_ROWS / 32; ++s) {
            acc = mfma16(*reinterpret_cast<const bf16x8 *>(At + ra + 32 * s),
                         *reinterpret_cast<const bf16x8 *>(Bt + rb + 32 * s), acc);
            acc = mfma16(*reinterpret_cast<const bf16x8 *>(Ap + ra + 32 * s),
                         *reinterpret_cast<const bf16x8 *>(Bp + rb + 32 * s), acc);
        }
        const int n = 16 * nt + col;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = 16 * mt + 4 * (lane >> 4) + r;
            if (w2) P[NH * KIN + m * NH + n] = acc[r];
            else P[m * KIN + n] = acc[r];
        }
    }
}

// Parameters: reduce the gradient parts in a fixed order, scale by
// -(2/n_comp) (the force-matching loss), then (mode 1) an Adam step of the
// FP32 master copy, or (mode 0) only the gradient into grad_out; mode 2
// only re-derives the device copies from theta.  Every mode rewrites the
// BF16 matrices (and transposes) and the FP32 vectors the kernels read.
// grid: ceil(2 NPE / 256), 256 threads.
struct NNDev {
    __bf16 *w1, *w1t, *w2, *w2t;
    float *b1, *b2, *w3, *b3;
};
struct AdamArgs {
    float lr, beta1, beta2, eps, c1, c2;   // c1 = 1 - beta1^t, c2 = 1 - beta2^t
};
__global__ __launch_bounds__(256) void k_nn_update(int mode, const float *__restrict__ part,
                                                   int nco, int nch, const float *__restrict__ tp,
                                                   int nto, int nth, double gscale,
                                                   float *__restrict__ theta,
                                                   float *__restrict__ am, float *__restrict__ av,
                                                   AdamArgs ad, float *__restrict__ grad_out,
                                                   NNDev D)
{
    const int pidx = blockIdx.x * 256 + threadIdx.x;
    if (pidx >= 2 * NPE) return;
    const int e = pidx / NPE, o = pidx - e * NPE;
    // which block of the packed element vector
    int kind, n = 0, k = 0;
    if (o < NH * NIN) { kind = 0; n = o / NIN; k = o - n * NIN; }
    else if (o < NH * NIN + NH) { kind = 1; n = o - NH * NIN; }
    else if (o < NH * NIN + NH + NH * NH) { kind = 2; const int q = o - NH * NIN - NH; n = q / NH; k = q - n * NH; }
    else if (o < NH * NIN + 2 * NH + NH * NH) { kind = 3; n = o - NH * NIN - NH - NH * NH; }
    else if (o < NPE - 1) { kind = 4; n = o - NH * NIN - 2 * NH - NH * NH; }
    else { kind = 5; }
    float th = theta[pidx];
    if (mode != 2) {
        double s = 0.0;
        if (kind == 0 || kind == 2) {
            const int c0 = e == 0 ? 0 : nco, c1 = e == 0 ? nco : nco + nch;
            const long off = kind == 0 ? (long)n * KIN + k : (long)NH * KIN + (long)n * NH + k;
            for (int c = c0; c < c1; ++c) s += (double)part[(long)c * WG_PART + off];
        } else if (kind != 5) {
            const int t0 = e == 0 ? 0 : nto, t1 = e == 0 ? nto : nto + nth;
            const int slot = kind == 1 ? 0 : (kind == 3 ? 1 : 2);
            for (int t = t0; t < t1; ++t) s += (double)tp[((long)t * 3 + slot) * NH + n];
        }
        const float gr = (float)(gscale * s);
        if (mode == 0) {
            grad_out[pidx] = gr;
        } else {
            const float m = ad.beta1 * am[pidx] + (1.0f - ad.beta1) * gr;
            const float v = ad.beta2 * av[pidx] + (1.0f - ad.beta2) * gr * gr;
            am[pidx] = m;
            av[pidx] = v;
            th -= ad.lr * (m / ad.c1) / (sqrtf(v / ad.c2) + ad.eps);
            theta[pidx] = th;
        }
    }
    switch (kind) {
    case 0:
        D.w1[(long)e * W1E + n * KIN + k] = (__bf16)th;
        D.w1t[(long)e * W1E + k * NH + n] = (__bf16)th;
        break;
    case 1: D.b1[e * NH + n] = th; break;
    case 2:
        D.w2[(long)e * W2E + n * NH + k] = (__bf16)th;
        D.w2t[(long)e * W2E + k * NH + n] = (__bf16)th;
        break;
    case 3: D.b2[e * NH + n] = th; break;
    case 4: D.w3[e * NH + n] = th; break;
    default: D.b3[e] = th; break;
    }
}

}  // namespace nn
}  // namespace qtp
