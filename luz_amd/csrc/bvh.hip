// bvh.hip — BLAS and TLAS builds on gfx950.
//
// Replaces the driver builds behind vkw::CmdBuildBLAS / vkw::CmdBuildTLAS
// (VulkanWrapper.cpp:1089-1104, 1106-1139) with a deterministic LBVH
// (Karras 2012, "Maximizing Parallelism in the Construction of BVHs,
// Octrees, and k-d Trees"):
//   1. padded triangle boxes + FP64 centroid bounds (block reductions);
//   2. 30-bit Morton codes in the FP64 frame;
//   3. stable LSD radix sort, 4 passes of 8 bits: per-tile histograms, one
//      scan, and a scatter that ranks keys inside a tile with wave ballots
//      (8 ballots give each lane the set of lanes holding the same digit);
//   4. Karras hierarchy, one thread per internal node;
//   5. triangle gather into slot order (36-byte records);
//   6. bottom-up refit, one thread per leaf, second arrival continues.
// The TLAS build runs in one workgroup (<= 8192 instances): world boxes,
// FP64 frame, Morton codes, bitonic sort of (code, index) in LDS, Karras,
// refit.  Both produce byte-identical nodes to a sequential CPU build.
#include "internal.hpp"

namespace luz {
namespace {

constexpr int kBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;

__device__ inline uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Morton code of the FP32 box centroid 0.5 (lo + hi) in the frame {origin[3],
// scale[3]}: the FP64 centroid-bounds reduction rounded to FP32 (morton_frame32).
// Quantisation is FP32; only the reduction that sets the frame runs in FP64.
__device__ inline uint32_t morton_code(const float* frame, const float* box) {
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c = 0.5f * (box[k] + box[3 + k]);
        float t = floorf((c - frame[k]) * frame[3 + k]);
        if (!(t >= 0.0f)) t = 0.0f;
        if (t > 1023.0f) t = 1023.0f;
        q[k] = (uint32_t)t;
    }
    return (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
}

// The FP32 quantisation frame from the FP64 centroid bounds mn, mx of one
// axis: origin mn and scale 1024 / (mx - mn) (0 for a flat axis), each
// rounded once to FP32.
__device__ inline void morton_frame32(double mn, double mx, float& origin, float& scale) {
    const double e = mx - mn;
    origin = (float)mn;
    scale = e > 0.0 ? (float)(1024.0 / e) : 0.0f;
}

// Padded triangle box (max |coord| * 2^-16 on every side).
__device__ inline void tri_box(V3 a, V3 b, V3 c, float* box) {
    float lo0 = fmin_(fmin_(a.x, b.x), c.x), hi0 = fmax_(fmax_(a.x, b.x), c.x);
    float lo1 = fmin_(fmin_(a.y, b.y), c.y), hi1 = fmax_(fmax_(a.y, b.y), c.y);
    float lo2 = fmin_(fmin_(a.z, b.z), c.z), hi2 = fmax_(fmax_(a.z, b.z), c.z);
    float m = 0.0f;
    m = fmax_(m, fabsf(lo0)); m = fmax_(m, fabsf(hi0));
    m = fmax_(m, fabsf(lo1)); m = fmax_(m, fabsf(hi1));
    m = fmax_(m, fabsf(lo2)); m = fmax_(m, fabsf(hi2));
    float pad = m * 0x1p-16f;
    box[0] = lo0 - pad; box[1] = lo1 - pad; box[2] = lo2 - pad;
    box[3] = hi0 + pad; box[4] = hi1 + pad; box[5] = hi2 + pad;
}

__device__ inline V3 load_pos(const uint8_t* base, uint32_t stride, uint32_t i) {
    const float* p = (const float*)(base + (size_t)i * stride);
    return v3(p[0], p[1], p[2]);
}

template <typename T, typename Op>
__device__ inline T block_reduce(T v, T* sh, Op op) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] = op(sh[tid], sh[tid + s]);
        __syncthreads();
    }
    T r = sh[0];
    __syncthreads();
    return r;
}

struct DMin { __device__ double operator()(double a, double b) const { return b < a ? b : a; } };
struct DMax { __device__ double operator()(double a, double b) const { return b > a ? b : a; } };

// Index of vertex k of triangle t, clamped into range; out-of-range
// indices set the error flag (the build then fails) instead of reading
// outside the vertex buffer.
__device__ inline uint32_t vidx(const uint32_t* idx, uint32_t i, uint32_t nv, uint32_t* error) {
    uint32_t v = idx[i];
    if (v >= nv) { atomicOr(error, 2u); v = 0; }
    return v;
}

__global__ __launch_bounds__(kBlock) void k_tri_boxes(const uint8_t* verts, uint32_t stride,
                                                      uint32_t nv, const uint32_t* idx, uint32_t n,
                                                      float* boxes, double* partial,
                                                      uint32_t* error) {
    __shared__ double sh[kBlock];
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    double c[3] = {INFINITY, INFINITY, INFINITY}, C[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (t < n) {
        V3 a = load_pos(verts, stride, vidx(idx, 3 * t, nv, error));
        V3 b = load_pos(verts, stride, vidx(idx, 3 * t + 1, nv, error));
        V3 cc = load_pos(verts, stride, vidx(idx, 3 * t + 2, nv, error));
        float bx[6];
        tri_box(a, b, cc, bx);
#pragma unroll
        for (int k = 0; k < 6; ++k) boxes[(size_t)t * 6 + k] = bx[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double ck = 0.5 * ((double)bx[k] + (double)bx[3 + k]);
            c[k] = ck; C[k] = ck;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double mn = block_reduce(c[k], sh, DMin());
        double mx = block_reduce(C[k], sh, DMax());
        if (threadIdx.x == 0) { partial[blockIdx.x * 6 + k] = mn; partial[blockIdx.x * 6 + 3 + k] = mx; }
    }
}

// frame = {cmin[3], scale[3]} with scale = 1024/extent (0 for a flat axis).
__global__ __launch_bounds__(kBlock) void k_frame(const double* partial, uint32_t nparts,
                                                  float* frame) {
    __shared__ double sh[kBlock];
    for (int k = 0; k < 3; ++k) {
        double mn = INFINITY, mx = -INFINITY;
        for (uint32_t i = threadIdx.x; i < nparts; i += kBlock) {
            double a = partial[i * 6 + k], b = partial[i * 6 + 3 + k];
            mn = a < mn ? a : mn;
            mx = b > mx ? b : mx;
        }
        mn = block_reduce(mn, sh, DMin());
        mx = block_reduce(mx, sh, DMax());
        if (threadIdx.x == 0) morton_frame32(mn, mx, frame[k], frame[3 + k]);
    }
}

__global__ __launch_bounds__(kBlock) void k_morton(const float* boxes, uint32_t n,
                                                   const float* frame, uint32_t* keys,
                                                   uint32_t* vals) {
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    keys[t] = morton_code(frame, boxes + (size_t)t * 6);
    vals[t] = t;
}

// ---------------------------------------------------------------- radix sort
__global__ __launch_bounds__(kBlock) void k_sort_hist(const uint32_t* keys, uint32_t n,
                                                      int shift, uint32_t* hist,
                                                      uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        uint32_t i = base + r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of count entries with one workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan(uint32_t* data, uint32_t count) {
    __shared__ uint32_t sh[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < count; base += 1024) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < count ? data[i] : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint32_t add = threadIdx.x >= (uint32_t)off ? sh[threadIdx.x - off] : 0u;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        uint32_t incl = sh[threadIdx.x];
        uint32_t c = carry;
        if (i < count) data[i] = c + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c + incl;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_sort_scatter(const uint32_t* keys_in,
                                                         const uint32_t* vals_in, uint32_t n,
                                                         int shift, const uint32_t* offsets,
                                                         uint32_t nblocks, uint32_t* keys_out,
                                                         uint32_t* vals_out) {
    __shared__ uint32_t digit_base[256];
    __shared__ uint32_t wave_cnt[kBlock / 64][256];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    digit_base[tid] = offsets[tid * nblocks + blockIdx.x];
    for (int w = 0; w < kBlock / 64; ++w) wave_cnt[w][tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kSortItems; ++r) {
        uint32_t i = base + r * kBlock + tid;
        bool valid = i < n;
        uint32_t key = valid ? keys_in[i] : 0u;
        uint32_t val = valid ? vals_in[i] : 0u;
        uint32_t d = (key >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wave_cnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t run = digit_base[tid];
            for (int w = 0; w < kBlock / 64; ++w) {
                uint32_t c = wave_cnt[w][tid];
                wave_cnt[w][tid] = run;
                run += c;
            }
            digit_base[tid] = run;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = wave_cnt[wave][d] + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        for (int w = 0; w < kBlock / 64; ++w) wave_cnt[w][tid] = 0;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- hierarchy
__device__ inline int delta_k(const uint32_t* k, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

// Karras split for internal node i; writes children and parents.
__device__ inline void karras_node(const uint32_t* keys, int n, int i, Node* nodes,
                                   uint32_t* leaf_parent, uint32_t* node_end) {
    int d = (delta_k(keys, n, i, i + 1) - delta_k(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta_k(keys, n, i, i - d);
    int lmax = 2;
    while (delta_k(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta_k(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta_k(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta_k(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    if (node_end) node_end[i] = (uint32_t)j;   // node i covers leaf slots lo..hi
    uint32_t left = (lo == gamma) ? (kLeaf | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t right = (hi == gamma + 1) ? (kLeaf | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    Node& nd = nodes[i];
    nd.child[0] = left;
    nd.child[1] = right;
    nd.pad = 0;
    if (i == 0) nd.parent = kNone;
    if (left & kLeaf) leaf_parent[left & ~kLeaf] = (uint32_t)i; else nodes[left].parent = (uint32_t)i;
    if (right & kLeaf) leaf_parent[right & ~kLeaf] = (uint32_t)i; else nodes[right].parent = (uint32_t)i;
}

__global__ __launch_bounds__(kBlock) void k_karras(const uint32_t* keys, uint32_t n, Node* nodes,
                                                   uint32_t* leaf_parent, uint32_t* node_end,
                                                   uint32_t* counters) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    karras_node(keys, (int)n, (int)i, nodes, leaf_parent, node_end);
    counters[i] = 0;
}

__global__ __launch_bounds__(kBlock) void k_gather(const uint8_t* verts, uint32_t stride,
                                                   uint32_t nv, const uint32_t* idx,
                                                   const uint32_t* order, uint32_t n,
                                                   const float* boxes, float* tris, float* nrms,
                                                   uint32_t* prims, uint32_t* slot_idx,
                                                   float* slot_box, uint32_t* error) {
    uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    uint32_t p = order[s];
    prims[s] = p;
    bool has_n = stride >= 24;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        uint32_t vi = vidx(idx, 3 * p + v, nv, error);
        slot_idx[3 * s + v] = vi;
        const float* q = (const float*)(verts + (size_t)vi * stride);
        tris[(size_t)s * 9 + v * 3 + 0] = q[0];
        tris[(size_t)s * 9 + v * 3 + 1] = q[1];
        tris[(size_t)s * 9 + v * 3 + 2] = q[2];
        nrms[(size_t)s * 9 + v * 3 + 0] = has_n ? q[3] : 0.0f;
        nrms[(size_t)s * 9 + v * 3 + 1] = has_n ? q[4] : 0.0f;
        nrms[(size_t)s * 9 + v * 3 + 2] = has_n ? q[5] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) slot_box[(size_t)s * 6 + k] = boxes[(size_t)p * 6 + k];
}

__device__ inline void store_child_box(Node* nd, int side, const float* b) {
    nd->lox[side] = b[0]; nd->loy[side] = b[1]; nd->loz[side] = b[2];
    nd->hix[side] = b[3]; nd->hiy[side] = b[4]; nd->hiz[side] = b[5];
}


constexpr int kRefitChunk = 256;    // leaf slots per workgroup of the chunk pass
constexpr int kSpanThreads = 512;   // threads per BLAS of the span pass

// Union of a node's two child boxes, child 0 first (the operand order of
// fmin_/fmax_ matters for signed zeros).
__device__ inline void child_union(int side, const float* mine, const float* other, float* out) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const float b0 = side == 0 ? mine[k] : other[k], b1 = side == 0 ? other[k] : mine[k];
        out[k] = k < 3 ? fmin_(b0, b1) : fmax_(b0, b1);
    }
}

__device__ inline uint32_t chunk_job_of(const RefitJob* jobs, uint32_t njobs, uint32_t c) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].chunk_offset <= c) lo = mid; else hi = mid;
    }
    return lo;
}

// Chunk pass of a refit (or of a build): one workgroup per 256 consecutive
// leaf slots of a BLAS.  Every internal node covers a contiguous range of
// slots (Karras), so a node whose range lies inside the chunk has both
// children completed by this workgroup: the second child to arrive forms the
// node's box, with arrival counters and child boxes in LDS.  A node whose
// parent's range leaves the chunk stores its box in the parent and hands the
// parent to the span pass through the BLAS's span list.  Also zeroes the span
// pass's arrival counters of the chunk's node ids.  jobs == nullptr: `single`.
template <bool FROM_VERTS>
__global__ __launch_bounds__(kRefitChunk) void k_refit_chunks(const RefitJob* jobs, uint32_t njobs,
                                                              RefitJob single, uint32_t* error) {
    __shared__ uint32_t scnt[kRefitChunk];
    __shared__ float sbox[kRefitChunk * 12];
    const RefitJob& J = jobs ? jobs[chunk_job_of(jobs, njobs, blockIdx.x)] : single;
    const uint32_t c0 = (blockIdx.x - J.chunk_offset) * kRefitChunk;
    const uint32_t c1 = (c0 + kRefitChunk < J.n ? c0 + kRefitChunk : J.n) - 1;   // last slot
    const uint32_t s = c0 + threadIdx.x;
    scnt[threadIdx.x] = 0;
    if (s < J.n - 1) J.counters[s] = 0;
    __syncthreads();
    if (s >= J.n) return;
    float box[6];
    if (FROM_VERTS) {
        // slot_idx was validated by the build
        V3 v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = load_pos(J.verts, J.stride, J.slot_idx[3 * s + k]);
            J.tris[(size_t)s * 9 + k * 3 + 0] = v[k].x;
            J.tris[(size_t)s * 9 + k * 3 + 1] = v[k].y;
            J.tris[(size_t)s * 9 + k * 3 + 2] = v[k].z;
        }
        tri_box(v[0], v[1], v[2], box);
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) box[k] = J.leaf_box[(size_t)s * 6 + k];
    }
    uint32_t c = kLeaf | s;
    uint32_t p = J.leaf_parent[s];
    for (int level = 0; p != kNone; ++level) {
        if (level > 80) { atomicOr(error, 1u); return; }
        Node* nd = &J.nodes[p];
        const int side = nd->child[0] == c ? 0 : 1;
        store_child_box(nd, side, box);
        const uint32_t e = J.node_end[p];
        const uint32_t lo = p < e ? p : e, hi = p < e ? e : p;
        if (lo < c0 || hi > c1) {
            const uint32_t k = atomicAdd(&J.span[0], 1u);
            J.span[1 + k] = p;
            return;
        }
        float* sb = &sbox[(p - c0) * 12];
#pragma unroll
        for (int k = 0; k < 6; ++k) sb[side * 6 + k] = box[k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (atomicAdd(&scnt[p - c0], 1u) == 0) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        float other[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) other[k] = sb[(1 - side) * 6 + k];
        float u[6];
        child_union(side, box, other, u);
#pragma unroll
        for (int k = 0; k < 6; ++k) box[k] = u[k];
        c = p;
        p = nd->parent;
    }
}

// Span pass: one workgroup per BLAS completes the nodes whose slot range
// crosses a chunk boundary, climbing from the span list (each entry is one
// arrival at a node whose child box is already stored in it).  Every arrival
// at these nodes comes from this workgroup, so arrival counters in global
// memory with workgroup-scope ordering suffice.  Leaves the list empty.
__global__ __launch_bounds__(kSpanThreads) void k_refit_spans(const RefitJob* jobs, RefitJob single,
                                                              uint32_t* error) {
    const RefitJob& J = jobs ? jobs[blockIdx.x] : single;
    const uint32_t count = J.span[0];
    for (uint32_t i = threadIdx.x; i < count; i += kSpanThreads) {
        uint32_t p = J.span[1 + i];
        for (int level = 0;; ++level) {
            if (level > 80) { atomicOr(error, 1u); break; }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (__hip_atomic_fetch_add(&J.counters[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
                break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const Node* nd = &J.nodes[p];
            const float b0[6] = {nd->lox[0], nd->loy[0], nd->loz[0], nd->hix[0], nd->hiy[0], nd->hiz[0]};
            const float b1[6] = {nd->lox[1], nd->loy[1], nd->loz[1], nd->hix[1], nd->hiy[1], nd->hiz[1]};
            float u[6];
            child_union(0, b0, b1, u);
            const uint32_t q = nd->parent;
            if (q == kNone) break;
            Node* nq = &J.nodes[q];
            store_child_box(nq, nq->child[0] == p ? 0 : 1, u);
            p = q;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) J.span[0] = 0;
}

// BLAS of 0 or 1 triangles: a single root with an empty second child.
__global__ void k_small_root(const uint8_t* verts, uint32_t stride, uint32_t nv,
                             const uint32_t* idx, uint32_t n, Node* nodes, float* tris,
                             float* nrms, uint32_t* prims, uint32_t* slot_idx, uint32_t* error) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Node nd;
    float empty[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    float b[6];
    if (n == 1) {
        V3 v[3];
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = vidx(idx, k, nv, error);
            slot_idx[k] = vi;
            const float* q = (const float*)(verts + (size_t)vi * stride);
            v[k] = v3(q[0], q[1], q[2]);
            tris[k * 3 + 0] = q[0]; tris[k * 3 + 1] = q[1]; tris[k * 3 + 2] = q[2];
            nrms[k * 3 + 0] = stride >= 24 ? q[3] : 0.0f;
            nrms[k * 3 + 1] = stride >= 24 ? q[4] : 0.0f;
            nrms[k * 3 + 2] = stride >= 24 ? q[5] : 0.0f;
        }
        tri_box(v[0], v[1], v[2], b);
        prims[0] = 0;
    } else {
        for (int k = 0; k < 6; ++k) b[k] = empty[k];
    }
    store_child_box(&nd, 0, b);
    store_child_box(&nd, 1, empty);
    nd.child[0] = n == 1 ? kLeaf : kEmpty;
    nd.child[1] = kEmpty;
    nd.parent = kNone;
    nd.pad = 0;
    nodes[0] = nd;
}

// ---------------------------------------------------------------- 8-wide form
// Collapse of the binary LBVH into 8-wide nodes, one breadth-first level per
// pass: k_wide_collapse lists each wide node's slots (repeatedly replacing the
// internal slot of largest surface area, first on ties, by its two children),
// two scans number the next level's nodes and this level's leaves, and
// k_wide_emit quantises the child boxes and writes nodes and leaf triangles.
struct WRef {
    uint32_t c;
    float lo[3], hi[3];
    uint32_t pad;
};

__device__ inline float surface_area(const float* lo, const float* hi) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

__device__ inline WRef child_ref(const Node& n, int c) {
    WRef r;
    r.c = n.child[c];
    r.lo[0] = n.lox[c]; r.lo[1] = n.loy[c]; r.lo[2] = n.loz[c];
    r.hi[0] = n.hix[c]; r.hi[1] = n.hiy[c]; r.hi[2] = n.hiz[c];
    r.pad = 0;
    return r;
}

__global__ __launch_bounds__(kBlock) void k_wide_collapse(const Node* bn, const uint32_t* level,
                                                          uint32_t count, WRef* slots_out,
                                                          uint32_t* nslots, uint32_t* n_int,
                                                          uint32_t* n_leaf) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    WRef s[8];
    int n = 0;
    const Node& root = bn[level[i]];
    for (int c = 0; c < 2; ++c)
        if (root.child[c] != kEmpty) s[n++] = child_ref(root, c);
    while (n < 8) {
        int best = -1;
        float best_sa = 0.0f;
        for (int k = 0; k < n; ++k) {
            if (s[k].c & kLeaf) continue;
            float sa = surface_area(s[k].lo, s[k].hi);
            if (best < 0 || sa > best_sa) { best = k; best_sa = sa; }
        }
        if (best < 0) break;
        const Node& in = bn[s[best].c];
        for (int k = n; k > best + 1; --k) s[k] = s[k - 1];
        s[best] = child_ref(in, 0);
        s[best + 1] = child_ref(in, 1);
        ++n;
    }
    uint32_t ni = 0, nl = 0;
    for (int k = 0; k < n; ++k) {
        slots_out[(size_t)i * 8 + k] = s[k];
        if (s[k].c & kLeaf) ++nl; else ++ni;
    }
    nslots[i] = (uint32_t)n;
    n_int[i] = ni;
    n_leaf[i] = nl;
}

__device__ inline void quant_frame(float lo, float hi, float& origin, int& e) {
    float t = (hi - lo) / 255.0f;
    int e0 = -126, em = -126;
    if (t > 0.0f) frexpf(t, &e0);
    float mag = fabsf(lo) > fabsf(hi) ? fabsf(lo) : fabsf(hi);
    if (mag > 0.0f) frexpf(mag, &em);
    e = e0 > em - 23 ? e0 : em - 23;
    if (e < -126) e = -126;
    for (;;) {
        float sc = ldexpf(1.0f, e);
        origin = floorf(lo / sc) * sc;
        if (origin > lo) origin = origin - sc;
        if (origin + 255.0f * sc >= hi) break;
        ++e;
    }
}

__device__ inline uint8_t quant_lo(float v, float origin, float sc) {
    float q = floorf((v - origin) / sc);
    q = q < 0.0f ? 0.0f : (q > 255.0f ? 255.0f : q);
    while (q > 0.0f && origin + q * sc > v) q = q - 1.0f;
    return (uint8_t)q;
}

__device__ inline uint8_t quant_hi(float v, float origin, float sc) {
    float q = ceilf((v - origin) / sc);
    q = q < 0.0f ? 0.0f : (q > 255.0f ? 255.0f : q);
    while (q < 255.0f && origin + q * sc < v) q = q + 1.0f;
    return (uint8_t)q;
}

__global__ __launch_bounds__(kBlock) void k_wide_emit(const WRef* slots, const uint32_t* nslots,
                                                      uint32_t count, uint32_t level_start,
                                                      uint32_t next_start, uint32_t tri_start,
                                                      const uint32_t* int_off,
                                                      const uint32_t* leaf_off, Node8* out,
                                                      uint32_t* next_level, const float* tris,
                                                      const uint32_t* prims, float* tris8,
                                                      uint32_t* prims8) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const uint32_t n = nslots[i];
    const WRef* s = slots + (size_t)i * 8;
    Node8 w;
    uint32_t* wz = reinterpret_cast<uint32_t*>(&w);
    for (int k = 0; k < 24; ++k) wz[k] = 0;
    w.nchild = (uint8_t)n;
    w.child_base = next_start + int_off[i];
    w.tri_base = tri_start + leaf_off[i];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) { lo[a] = fmin_(lo[a], s[k].lo[a]); hi[a] = fmax_(hi[a], s[k].hi[a]); }
    float origin[3] = {0.0f, 0.0f, 0.0f};
    int e[3] = {0, 0, 0};
    if (n > 0)
        for (int a = 0; a < 3; ++a) quant_frame(lo[a], hi[a], origin[a], e[a]);
    w.ox = origin[0]; w.oy = origin[1]; w.oz = origin[2];
    w.ex = (int8_t)e[0]; w.ey = (int8_t)e[1]; w.ez = (int8_t)e[2];
    uint32_t ni = 0, nl = 0;
    for (uint32_t k = 0; k < n; ++k) {
        for (int a = 0; a < 3; ++a) {
            float sc = ldexpf(1.0f, e[a]);
            w.qlo[a][k] = quant_lo(s[k].lo[a], origin[a], sc);
            w.qhi[a][k] = quant_hi(s[k].hi[a], origin[a], sc);
        }
        if (s[k].c & kLeaf) {
            uint32_t slot = s[k].c & ~kLeaf;
            uint32_t dst = w.tri_base + nl;
            for (int q = 0; q < 9; ++q) tris8[(size_t)dst * 9 + q] = tris[(size_t)slot * 9 + q];
            prims8[dst] = prims[slot];
            w.meta[k] = (uint8_t)nl++;
        } else {
            next_level[int_off[i] + ni] = s[k].c;
            w.meta[k] = (uint8_t)(0x80u | ni++);
        }
    }
    out[level_start + i] = w;
}

// ---------------------------------------------------------------- TLAS
constexpr int kTlasThreads = 1024;
constexpr int kPlocRadius = 16;                              // clustering search radius (slots)
constexpr int kPlocTile = kTlasThreads + 2 * kPlocRadius;    // staged boxes per neighbour pass

// Threads of a TLAS build launch for m (a power of two >= the instance count):
// four waves up to 64 instances (the neighbour search then gives each cluster
// four threads, one per SIMD, and the barriers stay cheap), else kTlasThreads.
inline uint32_t tlas_threads(uint32_t m) { return m <= 64 ? 256u : (uint32_t)kTlasThreads; }

// Merge cost of two cluster boxes as the tile holds them ((lo.xyz, hi.x) and
// (hi.y, hi.z)): half the surface area of the union, FP32 as written, NaN as
// +inf.
__device__ inline float merge_cost(float4 a, float2 a2, float4 b, float2 b2) {
    const float lx = fmin_(a.x, b.x), ly = fmin_(a.y, b.y), lz = fmin_(a.z, b.z);
    const float hx = fmax_(a.w, b.w), hy = fmax_(a2.x, b2.x), hz = fmax_(a2.y, b2.y);
    const float ex = hx - lx, ey = hy - ly, ez = hz - lz;
    const float c = ex * ey + ey * ez + ez * ex;
    return c == c ? c : INFINITY;
}

// The LDS tile of cluster boxes: slot t holds (lo.x, lo.y, lo.z, hi.x) in
// a[t] and (hi.y, hi.z) in b[t], 24 bytes read by two loads.
struct BoxTile {
    float4* a;
    float2* b;
};
__device__ inline void tile_put(BoxTile T, int t, const float* box) {
    T.a[t] = make_float4(box[0], box[1], box[2], box[3]);
    T.b[t] = make_float2(box[4], box[5]);
}
__device__ inline void tile_get(BoxTile T, int t, float* box) {
    const float4 a = T.a[t];
    const float2 b = T.b[t];
    box[0] = a.x; box[1] = a.y; box[2] = a.z;
    box[3] = a.w; box[4] = b.x; box[5] = b.y;
}

// Part q (< 4) of the neighbour search of cluster i among clusters
// i - kPlocRadius .. i + kPlocRadius (within 0 .. count - 1), from a tile that
// holds cluster c at slot c - base + kPlocRadius: candidates 8q .. 8q + 7 of
// the 32 in index order, their boxes read from LDS together (slots outside
// 0 .. count - 1 are read, as the tile spans them, and skipped).  Keeps in
// (best, bj) the least merge cost, the smaller index on equal cost.
__device__ inline void ploc_part(BoxTile tile, uint32_t i, uint32_t base, uint32_t count, int q, float& best,
                                 uint32_t& bj) {
    const int ti = (int)(i - base) + kPlocRadius;
    const float4 a = tile.a[ti];
    const float2 a2 = tile.b[ti];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int k = 8 * q + u, d = k < kPlocRadius ? k - kPlocRadius : k - kPlocRadius + 1;
        const int j = (int)i + d;
        const float c = merge_cost(a, a2, tile.a[ti + d], tile.b[ti + d]);
        if (j >= 0 && j < (int)count && (bj == kNone || c < best)) { best = c; bj = (uint32_t)j; }
    }
}

// Nearest neighbour of cluster i: least merge cost, the smaller index on
// equal cost.
__device__ inline uint32_t ploc_nearest(BoxTile tile, uint32_t i, uint32_t base, uint32_t count) {
    float best = INFINITY;
    uint32_t bj = kNone;
    for (int q = 0; q < 4; ++q) ploc_part(tile, i, base, count, q, best, bj);
    return bj;
}

// The same search by four consecutive threads (a quad), thread q taking part
// q; the quad's parts combine as (cost, index) pairs, least first, which is
// the order of the one-thread search.  All threads of the wave call it; the
// result is valid in every thread of quad i (i < count).
__device__ inline uint32_t ploc_nearest_quad(BoxTile tile, uint32_t i, uint32_t count) {
    float best = INFINITY;
    uint32_t bj = kNone;
    if (i < count) ploc_part(tile, i, 0, count, (int)(threadIdx.x & 3u), best, bj);
    for (int o = 1; o <= 2; o <<= 1) {
        const float ob = __shfl_xor(best, o);
        const uint32_t oj = __shfl_xor(bj, o);
        if (ob < best || (ob == best && oj < bj)) { best = ob; bj = oj; }
    }
    return bj;
}

// Workgroup barrier ordering LDS accesses only: global stores still in
// flight are not waited for.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Exclusive ranks of this thread's head and keep flags among the workgroup's
// threads (thread order), and their totals.  All threads must call it.
// LDS_ONLY: its barrier orders LDS accesses only.
template <bool LDS_ONLY = false>
__device__ inline void block_ranks(bool head, bool keep, uint32_t (*s_wave)[kTlasThreads / 64],
                                   uint32_t& hr, uint32_t& kr, uint32_t& th, uint32_t& tk) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint64_t bh = __ballot(head), bk = __ballot(keep);
    if (lane == 0) {
        s_wave[0][wave] = (uint32_t)__popcll(bh);
        s_wave[1][wave] = (uint32_t)__popcll(bk);
    }
    if (LDS_ONLY) lds_barrier();
    else __syncthreads();
    uint32_t oh = 0, ok = 0;
    th = 0;
    tk = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
        const uint32_t ch = s_wave[0][w], ck = s_wave[1][w];
        if (w < wave) { oh += ch; ok += ck; }
        th += ch;
        tk += ck;
    }
    hr = oh + (uint32_t)__popcll(bh & below);
    kr = ok + (uint32_t)__popcll(bk & below);
}

// Node idx joining clusters a (child 0) and b (child 1); the children that are
// nodes get idx as their parent.
__device__ inline void ploc_node(Node* nodes, uint32_t idx, const float* a, const float* b, uint32_t ra,
                                 uint32_t rb) {
    Node nd;
    store_child_box(&nd, 0, a);
    store_child_box(&nd, 1, b);
    nd.child[0] = ra;
    nd.child[1] = rb;
    nd.parent = kNone;
    nd.pad = 0;
    nodes[idx] = nd;
    if (!(ra & kLeaf)) nodes[ra].parent = idx;
    if (!(rb & kLeaf)) nodes[rb].parent = idx;
}

// The same node written without its parent field, which the merge that takes
// the node in sets (kNone for the root, node 0): every byte of a node is then
// stored once, so rounds that create and link nodes need not wait for each
// other's stores.
__device__ inline void ploc_node_unlinked(Node* nodes, uint32_t idx, const float* a, const float* b,
                                          uint32_t ra, uint32_t rb) {
    Node* nd = nodes + idx;
    float4* q = reinterpret_cast<float4*>(nd);
    q[0] = make_float4(a[0], b[0], a[3], b[3]);   // lox[2], hix[2]
    q[1] = make_float4(a[1], b[1], a[4], b[4]);   // loy[2], hiy[2]
    q[2] = make_float4(a[2], b[2], a[5], b[5]);   // loz[2], hiz[2]
    *reinterpret_cast<uint2*>(nd->child) = make_uint2(ra, rb);
    nd->pad = 0;
    if (idx == 0) nd->parent = kNone;
    if (!(ra & kLeaf)) nodes[ra].parent = idx;
    if (!(rb & kLeaf)) nodes[rb].parent = idx;
}

__device__ inline void tlas_build_body(const TlasJob& J, uint32_t* error) {
    const InstIn* in = J.in;
    const uint32_t n = J.n, m = J.m;
    float* wbox = J.wbox;
    Node* nodes = J.nodes;
    InstRec* recs = J.recs;
    InstRec* recs_w = J.recs_w;
    InstAux* aux = J.aux;
    float* w2o_out = J.w2o_out;
    uint32_t* order_out = J.order_out;
    // tlas_lds_bytes(m) bytes, m a power of two >= n: the sort keys, then the
    // clustering's box tile, neighbour and reference arrays in the same memory
    // (within 64 KB for 8192 instances).
    extern __shared__ __align__(16) uint64_t skeys[];
    const int tid = threadIdx.x;
    // launched with 64 threads (one wave, cheap barriers) up to 64 instances,
    // else with kTlasThreads
    const uint32_t nt = blockDim.x;

    // 1. world boxes of the BLAS root boxes, scene magnitude
    float mloc = 0.0f;
    for (uint32_t i = tid; i < n; i += nt) {
        const Node& r = in[i].nodes[0];
        float rlo[3] = {fmin_(r.lox[0], r.lox[1]), fmin_(r.loy[0], r.loy[1]), fmin_(r.loz[0], r.loz[1])};
        float rhi[3] = {fmax_(r.hix[0], r.hix[1]), fmax_(r.hiy[0], r.hiy[1]), fmax_(r.hiz[0], r.hiz[1])};
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = 0; c < 8; ++c) {
            V3 p = v3((c & 1) ? rhi[0] : rlo[0], (c & 2) ? rhi[1] : rlo[1], (c & 4) ? rhi[2] : rlo[2]);
            V3 w = xform_point(in[i].o2w, p);
            lo[0] = fmin_(lo[0], w.x); hi[0] = fmax_(hi[0], w.x);
            lo[1] = fmin_(lo[1], w.y); hi[1] = fmax_(hi[1], w.y);
            lo[2] = fmin_(lo[2], w.z); hi[2] = fmax_(hi[2], w.z);
        }
        for (int k = 0; k < 3; ++k) {
            mloc = fmax_(mloc, fabsf(lo[k]));
            mloc = fmax_(mloc, fabsf(hi[k]));
            wbox[i * 6 + k] = lo[k];
            wbox[i * 6 + 3 + k] = hi[k];
        }
    }
    // scene magnitude: wave maxima, then the 16 wave values (min/max are exact,
    // so the order of the reduction does not change the result)
    __shared__ float s_red[kTlasThreads / 64];
    __shared__ double s_redd[6][kTlasThreads / 64];
    const uint32_t wave = (uint32_t)tid >> 6, lane = (uint32_t)tid & 63u;
    for (int o = 32; o > 0; o >>= 1) mloc = fmax_(mloc, __shfl_xor(mloc, o));
    if (lane == 0) s_red[wave] = mloc;
    __syncthreads();
    float m_all = 0.0f;
    for (uint32_t w = 0; w < nt / 64; ++w) m_all = fmax_(m_all, s_red[w]);
    const float pad = m_all * 0x1p-14f;
    // 2. padded boxes, FP64 centroid frame
    double cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = tid; i < n; i += nt) {
        for (int k = 0; k < 3; ++k) {
            float lo = wbox[i * 6 + k] - pad, hi = wbox[i * 6 + 3 + k] + pad;
            wbox[i * 6 + k] = lo;
            wbox[i * 6 + 3 + k] = hi;
            double c = 0.5 * ((double)lo + (double)hi);
            cmn[k] = c < cmn[k] ? c : cmn[k];
            cmx[k] = c > cmx[k] ? c : cmx[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        for (int o = 32; o > 0; o >>= 1) {
            const double a = __shfl_xor(cmn[k], o), b = __shfl_xor(cmx[k], o);
            cmn[k] = a < cmn[k] ? a : cmn[k];
            cmx[k] = b > cmx[k] ? b : cmx[k];
        }
        if (lane == 0) {
            s_redd[k][wave] = cmn[k];
            s_redd[3 + k][wave] = cmx[k];
        }
    }
    __syncthreads();
    float s_frame[6];
    for (int k = 0; k < 3; ++k) {
        double mn = INFINITY, mx = -INFINITY;
        for (uint32_t w = 0; w < nt / 64; ++w) {
            mn = s_redd[k][w] < mn ? s_redd[k][w] : mn;
            mx = s_redd[3 + k][w] > mx ? s_redd[3 + k][w] : mx;
        }
        morton_frame32(mn, mx, s_frame[k], s_frame[3 + k]);
    }
    // 3. Morton codes, bitonic sort of (code << 32 | index)
    auto sort_key = [&](uint32_t i) -> uint64_t {
        if (i >= n) return ~0ull;
        return ((uint64_t)morton_code(s_frame, wbox + i * 6) << 32) | i;
    };
    if (m <= 64) {
        // one wave sorts in registers: lane i holds key i, partners by xor
        if (tid < 64) {
            uint64_t key = sort_key((uint32_t)tid);
            for (uint32_t k = 2; k <= m; k <<= 1) {
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    const uint32_t lo = __shfl_xor((uint32_t)key, (int)j);
                    const uint32_t hi = __shfl_xor((uint32_t)(key >> 32), (int)j);
                    const uint64_t other = ((uint64_t)hi << 32) | lo;
                    const bool lower = ((uint32_t)tid & j) == 0, up = ((uint32_t)tid & k) == 0;
                    const uint64_t mn = other < key ? other : key, mx = other < key ? key : other;
                    key = lower == up ? mn : mx;
                }
            }
            if ((uint32_t)tid < m) skeys[tid] = key;
        }
        __syncthreads();
    } else {
        for (uint32_t i = tid; i < m; i += nt) skeys[i] = sort_key(i);
        __syncthreads();
    }
    for (uint32_t k = 2; m > 64 && k <= m; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < m; i += nt) {
                uint32_t ixj = i ^ j;
                if (ixj > i) {
                    uint64_t a = skeys[i], b = skeys[ixj];
                    bool up = (i & k) == 0;
                    if ((a > b) == up) { skeys[i] = b; skeys[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
    // 4. leaf records in slot order; the clusters start as the leaves
    float* cb = J.cbox[0];
    float* cb2 = J.cbox[1];
    uint32_t* cr = J.cref[0];
    uint32_t* cr2 = J.cref[1];
    for (uint32_t s = tid; s < n; s += nt) {
        uint32_t i = (uint32_t)(skeys[s] & 0xFFFFFFFFu);
        const InstIn& src = in[i];
        InstRec rec;
        for (int k = 0; k < 12; ++k) rec.w2o[k] = src.w2o[k];
        rec.nodes = src.nodes;
        rec.tris = src.tris;
        recs[s] = rec;
        if (recs_w) {
            rec.nodes = reinterpret_cast<const Node*>(src.wnodes);
            rec.tris = src.wtris;
            recs_w[s] = rec;
        }
        InstAux ax;
        for (int k = 0; k < 9; ++k) ax.nmat[k] = src.nmat[k];
        ax.custom_index = src.custom_index;
        ax.input_index = src.input_index;
        ax.pad = 0;
        ax.nrms = src.nrms;
        ax.prims = src.prims;
        aux[s] = ax;
        for (int k = 0; k < 12; ++k) w2o_out[s * 12 + k] = src.w2o[k];
        order_out[s] = src.input_index;
        for (int k = 0; k < 6; ++k) cb[s * 6 + k] = wbox[i * 6 + k];
        for (int k = 0; k < 6; ++k) J.slot_box[s * 6 + k] = wbox[i * 6 + k];
        cr[s] = kLeaf | s;
    }
    __syncthreads();   // the sort keys are dead from here on
    if (n >= 2) {
        // 5. locally-ordered clustering (PLOC, Meister & Bittner 2018): each
        // round every cluster finds the neighbour within kPlocRadius slots of
        // least merge cost, mutual pairs merge into a new node at the lower
        // position, and the survivors are compacted in order.  New nodes are
        // numbered downwards from n - 2 in merge order, so the root is node 0.
        BoxTile tile{reinterpret_cast<float4*>(skeys), nullptr};   // kPlocTile boxes
        tile.b = reinterpret_cast<float2*>(tile.a + kPlocTile);
        uint32_t* snn = reinterpret_cast<uint32_t*>(tile.b + kPlocTile);   // m entries
        uint32_t* sref = snn + m;                                              // kTlasThreads entries
        __shared__ uint32_t s_wave[2][kTlasThreads / 64];
        uint32_t count = n, next = n - 1;
        bool failed = false;
        // rounds over more clusters than threads: cluster arrays in global memory
        while (count > nt && !failed) {
            for (uint32_t base = 0; base < count; base += nt) {
                for (int t = tid; t < (int)nt + 2 * kPlocRadius; t += (int)nt) {
                    const int c = (int)base - kPlocRadius + t;
                    if (c >= 0 && c < (int)count) tile_put(tile, t, cb + (size_t)c * 6);
                }
                __syncthreads();
                const uint32_t i = base + tid;
                if (i < count) snn[i] = ploc_nearest(tile, i, base, count);
                __syncthreads();
            }
            uint32_t heads = 0, kept = 0;
            for (uint32_t base = 0; base < count; base += nt) {
                const uint32_t i = base + tid;
                bool head = false, keep = false;
                uint32_t j = 0;
                if (i < count) {
                    j = snn[i];
                    const bool mutual = snn[j] == i;
                    head = mutual && i < j;
                    keep = !(mutual && i > j);
                }
                uint32_t hr, kr, th, tk;
                block_ranks(head, keep, s_wave, hr, kr, th, tk);
                hr += heads;
                kr += kept;
                if (head) {
                    const uint32_t idx = next - 1 - hr;
                    ploc_node(nodes, idx, cb + (size_t)i * 6, cb + (size_t)j * 6, cr[i], cr[j]);
                    cr2[kr] = idx;
                    for (int k = 0; k < 3; ++k) {
                        cb2[kr * 6 + k] = fmin_(cb[i * 6 + k], cb[j * 6 + k]);
                        cb2[kr * 6 + 3 + k] = fmax_(cb[i * 6 + 3 + k], cb[j * 6 + 3 + k]);
                    }
                } else if (keep) {
                    cr2[kr] = cr[i];
                    for (int k = 0; k < 6; ++k) cb2[kr * 6 + k] = cb[i * 6 + k];
                }
                heads += th;
                kept += tk;
                __syncthreads();
            }
            failed = heads == 0;   // cannot happen: the least-cost pair is mutual
            count = kept;
            next -= heads;
            float* tb = cb; cb = cb2; cb2 = tb;
            uint32_t* tr = cr; cr = cr2; cr2 = tr;
        }
        // rounds over at most one cluster per thread: cluster i lives in thread
        // i's registers and in the LDS tile (box at i + kPlocRadius) and sref;
        // the nodes these rounds create are only stored, so their barriers
        // order LDS accesses alone
        float box[6];
        uint32_t ref = 0;
        if (tid < count) {
            for (int k = 0; k < 6; ++k) box[k] = cb[tid * 6 + k];
            tile_put(tile, tid + kPlocRadius, box);
            ref = cr[tid];
            sref[tid] = ref;
        }
        __syncthreads();
        while (count > 1 && !failed) {
            if (4 * count <= nt) {   // four threads per cluster
                const uint32_t i = (uint32_t)tid >> 2, b = ploc_nearest_quad(tile, i, count);
                if (i < count && (tid & 3) == 0) snn[i] = b;
            } else if (tid < count) {
                snn[tid] = ploc_nearest(tile, tid, 0, count);
            }
            lds_barrier();
            bool head = false, keep = false;
            uint32_t j = 0;
            if (tid < count) {
                j = snn[tid];
                const bool mutual = snn[j] == tid;
                head = mutual && tid < j;
                keep = !(mutual && tid > j);
            }
            uint32_t hr, kr, heads, kept;
            block_ranks<true>(head, keep, s_wave, hr, kr, heads, kept);
            if (head) {
                const uint32_t idx = next - 1 - hr;
                float other[6];
                tile_get(tile, j + kPlocRadius, other);
                ploc_node_unlinked(nodes, idx, box, other, ref, sref[j]);
                for (int k = 0; k < 3; ++k) {
                    box[k] = fmin_(box[k], other[k]);
                    box[3 + k] = fmax_(box[3 + k], other[3 + k]);
                }
                ref = idx;
            }
            lds_barrier();   // every read of this round's tile and sref is done
            if (keep) {
                tile_put(tile, kr + kPlocRadius, box);
                sref[kr] = ref;
            }
            lds_barrier();
            failed = heads == 0;
            count = kept;
            next -= heads;
            if (tid < count) {
                tile_get(tile, tid + kPlocRadius, box);
                ref = sref[tid];
            }
        }
        if (failed && tid == 0) atomicOr(error, 1u);
    } else if (tid == 0) {
        Node nd;
        float empty[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        if (n == 1) store_child_box(&nd, 0, wbox);
        else store_child_box(&nd, 0, empty);
        store_child_box(&nd, 1, empty);
        nd.child[0] = n == 1 ? kLeaf : kEmpty;
        nd.child[1] = kEmpty;
        nd.parent = kNone;
        nd.pad = 0;
        nodes[0] = nd;
    }
}

__global__ __launch_bounds__(kTlasThreads) void k_tlas_build(TlasJob J, uint32_t* error) {
    tlas_build_body(J, error);
}

// Many TLASes in one launch: workgroup i builds jobs[i].
__global__ __launch_bounds__(kTlasThreads) void k_tlas_build_batched(const TlasJob* jobs,
                                                                     uint32_t* error) {
    tlas_build_body(jobs[blockIdx.x], error);
}

// Batched refit: global triangle t belongs to the job whose tri_offset range holds it.
inline uint32_t cdiv(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

// Device flag set by build kernels on bad input or a malformed tree; read by
// luzrt_get_timings / the build call's own check.
uint32_t* build_error_flag(luzrt_ctx* ctx) {
    if (ensure(ctx, ctx->counters, kCounterBytes) != LUZRT_OK) return nullptr;
    return (uint32_t*)ctx->counters.ptr + kCntBuildError;
}

luzrt_status blas_build_device(luzrt_ctx* ctx, luzrt_blas* b, const void* vertices,
                               const uint32_t* indices) {
    ctx->release_clean = false;   // BLAS data written on the context's stream
    hipStream_t st = ctx->stream;
    uint32_t* error = build_error_flag(ctx);
    if (!error) return fail(ctx, LUZRT_E_OUT_OF_MEMORY, "no device memory for the error flag");
    const uint32_t n = b->tri_count;
    const uint8_t* verts = (const uint8_t*)vertices;
    if (n <= 1) {
        k_small_root<<<1, 64, 0, st>>>(verts, b->stride, b->vertex_count, indices, n, b->nodes,
                                       b->tris, b->nrms, b->prims, b->slot_idx, error);
        LUZ_HIP(ctx, hipGetLastError());
        return LUZRT_OK;
    }
    const uint32_t nb = cdiv(n, kBlock);
    const uint32_t sb = cdiv(n, kSortTile);
    // 10 sub-allocations, each rounded up to 256 bytes
    size_t bytes = (size_t)n * 24 + (size_t)nb * 48 + 64 + (size_t)n * 16 + (size_t)256 * sb * 4 +
                   (size_t)n * 24 + 10 * 256;
    uint8_t* scratch = nullptr;
    LUZ_HIP(ctx, hipMallocAsync((void**)&scratch, bytes, st));
    uint8_t* p = scratch;
    auto take = [&](size_t sz) { uint8_t* r = p; p += (sz + 255) & ~(size_t)255; return r; };
    float* boxes = (float*)take((size_t)n * 24);
    double* partial = (double*)take((size_t)nb * 48);
    float* frame = (float*)take(64);
    uint32_t* k0 = (uint32_t*)take((size_t)n * 4);
    uint32_t* v0 = (uint32_t*)take((size_t)n * 4);
    uint32_t* k1 = (uint32_t*)take((size_t)n * 4);
    uint32_t* v1 = (uint32_t*)take((size_t)n * 4);
    uint32_t* hist = (uint32_t*)take((size_t)256 * sb * 4);
    float* slot_box = (float*)take((size_t)n * 24);
    (void)bytes;

    k_tri_boxes<<<nb, kBlock, 0, st>>>(verts, b->stride, b->vertex_count, indices, n, boxes, partial,
                                       error);
    k_frame<<<1, kBlock, 0, st>>>(partial, nb, frame);
    k_morton<<<nb, kBlock, 0, st>>>(boxes, n, frame, k0, v0);
    uint32_t *kin = k0, *vin = v0, *kout = k1, *vout = v1;
    for (int pass = 0; pass < 4; ++pass) {
        int shift = pass * 8;
        k_sort_hist<<<sb, kBlock, 0, st>>>(kin, n, shift, hist, sb);
        k_scan<<<1, 1024, 0, st>>>(hist, 256 * sb);
        k_sort_scatter<<<sb, kBlock, 0, st>>>(kin, vin, n, shift, hist, sb, kout, vout);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    k_karras<<<cdiv(n - 1, kBlock), kBlock, 0, st>>>(kin, n, b->nodes, b->leaf_parent, b->node_end,
                                                      b->counters);
    k_gather<<<nb, kBlock, 0, st>>>(verts, b->stride, b->vertex_count, indices, vin, n, boxes,
                                    b->tris, b->nrms, b->prims, b->slot_idx, slot_box, error);
    const RefitJob J = refit_job(b, nullptr, slot_box);
    k_refit_chunks<false><<<refit_chunks(n), kRefitChunk, 0, st>>>(nullptr, 1, J, error);
    k_refit_spans<<<1, kSpanThreads, 0, st>>>(nullptr, J, error);
    LUZ_HIP(ctx, hipGetLastError());
    LUZ_HIP(ctx, hipFreeAsync(scratch, st));
    return LUZRT_OK;
}

luzrt_status blas_refit_device(luzrt_ctx* ctx, luzrt_blas* b, const void* vertices) {
    ctx->release_clean = false;   // BLAS data written on the context's stream
    hipStream_t st = ctx->stream;
    uint32_t* error = build_error_flag(ctx);
    if (!error) return fail(ctx, LUZRT_E_OUT_OF_MEMORY, "no device memory for the error flag");
    const uint32_t n = b->tri_count;
    if (n <= 1) {
        // rebuild the single root (indices are kept in slot order)
        k_small_root<<<1, 64, 0, st>>>((const uint8_t*)vertices, b->stride, b->vertex_count,
                                       b->slot_idx, n, b->nodes, b->tris, b->nrms, b->prims,
                                       b->slot_idx, error);
        LUZ_HIP(ctx, hipGetLastError());
        return LUZRT_OK;
    }
    const RefitJob J = refit_job(b, vertices, nullptr);
    k_refit_chunks<true><<<refit_chunks(n), kRefitChunk, 0, st>>>(nullptr, 1, J, error);
    k_refit_spans<<<1, kSpanThreads, 0, st>>>(nullptr, J, error);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

luzrt_status blas_wide_device(luzrt_ctx* ctx, luzrt_blas* b) {
    ctx->release_clean = false;   // BLAS data written on the context's stream
    hipStream_t st = ctx->stream;
    const uint32_t n = b->tri_count;
    const size_t cap = n > 1 ? n - 1 : 1;   // wide nodes <= binary internal nodes
    if (!b->wide) {
        LUZ_HIP(ctx, hipMalloc((void**)&b->wide, cap * sizeof(Node8)));
        LUZ_HIP(ctx, hipMalloc((void**)&b->tris8, (size_t)(n ? n : 1) * 36));
        LUZ_HIP(ctx, hipMalloc((void**)&b->prims8, (size_t)(n ? n : 1) * 4));
    }
    // scratch: two level lists, slots, counts and offsets
    // 8 sub-allocations, each rounded up to 256 bytes
    size_t bytes = cap * 4 * 2 + cap * 8 * sizeof(WRef) + cap * 4 * 5 + 8 * 256;
    uint8_t* scratch = nullptr;
    LUZ_HIP(ctx, hipMallocAsync((void**)&scratch, bytes, st));
    uint8_t* p = scratch;
    auto take = [&](size_t sz) { uint8_t* r = p; p += (sz + 255) & ~(size_t)255; return r; };
    uint32_t* lvl[2] = {(uint32_t*)take(cap * 4), (uint32_t*)take(cap * 4)};
    WRef* slots = (WRef*)take(cap * 8 * sizeof(WRef));
    uint32_t* nslots = (uint32_t*)take(cap * 4);
    uint32_t* n_int = (uint32_t*)take(cap * 4);
    uint32_t* n_leaf = (uint32_t*)take(cap * 4);
    uint32_t* int_off = (uint32_t*)take(cap * 4);
    uint32_t* leaf_off = (uint32_t*)take(cap * 4);
    (void)bytes;
    LUZ_HIP(ctx, hipMemsetAsync(lvl[0], 0, 4, st));   // level 0 = binary root
    uint32_t count = 1, level_start = 0, tri_start = 0;
    int cur = 0;
    while (count > 0) {
        const uint32_t blocks = cdiv(count, kBlock);
        k_wide_collapse<<<blocks, kBlock, 0, st>>>(b->nodes, lvl[cur], count, slots, nslots, n_int, n_leaf);
        LUZ_HIP(ctx, hipMemcpyAsync(int_off, n_int, (size_t)count * 4, hipMemcpyDeviceToDevice, st));
        LUZ_HIP(ctx, hipMemcpyAsync(leaf_off, n_leaf, (size_t)count * 4, hipMemcpyDeviceToDevice, st));
        k_scan<<<1, 1024, 0, st>>>(int_off, count);
        k_scan<<<1, 1024, 0, st>>>(leaf_off, count);
        uint32_t tail[4];
        LUZ_HIP(ctx, hipMemcpyAsync(&tail[0], int_off + count - 1, 4, hipMemcpyDeviceToHost, st));
        LUZ_HIP(ctx, hipMemcpyAsync(&tail[1], n_int + count - 1, 4, hipMemcpyDeviceToHost, st));
        LUZ_HIP(ctx, hipMemcpyAsync(&tail[2], leaf_off + count - 1, 4, hipMemcpyDeviceToHost, st));
        LUZ_HIP(ctx, hipMemcpyAsync(&tail[3], n_leaf + count - 1, 4, hipMemcpyDeviceToHost, st));
        LUZ_HIP(ctx, hipStreamSynchronize(st));
        const uint32_t next_count = tail[0] + tail[1], leaves = tail[2] + tail[3];
        const uint32_t next_start = level_start + count;
        if ((size_t)next_start + next_count > cap || (size_t)tri_start + leaves > (n ? n : 1)) {
            hipFreeAsync(scratch, st);
            return fail(ctx, LUZRT_E_DEVICE, "8-wide collapse overflowed its arrays");
        }
        k_wide_emit<<<blocks, kBlock, 0, st>>>(slots, nslots, count, level_start, next_start, tri_start,
                                               int_off, leaf_off, b->wide, lvl[cur ^ 1], b->tris,
                                               b->prims, b->tris8, b->prims8);
        LUZ_HIP(ctx, hipGetLastError());
        level_start = next_start;
        tri_start += leaves;
        count = next_count;
        cur ^= 1;
    }
    LUZ_HIP(ctx, hipFreeAsync(scratch, st));
    b->wide_count = level_start;
    b->wide_valid = true;
    return LUZRT_OK;
}

luzrt_status launch_tlas_batched(luzrt_ctx* ctx, const TlasJob* dev_jobs, uint32_t njobs,
                                 uint32_t max_m) {
    uint32_t* error = build_error_flag(ctx);
    if (!error) return fail(ctx, LUZRT_E_OUT_OF_MEMORY, "no device memory for the error flag");
    const size_t lds = tlas_lds_bytes(max_m);
    k_tlas_build_batched<<<njobs, tlas_threads(max_m), lds, ctx->stream>>>(dev_jobs, error);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

uint32_t refit_chunks(uint32_t tris) { return cdiv(tris, kRefitChunk); }

RefitJob refit_job(luzrt_blas* b, const void* vertices, const float* leaf_box) {
    RefitJob J;
    J.nodes = b->nodes;
    J.leaf_parent = b->leaf_parent;
    J.counters = b->counters;
    J.node_end = b->node_end;
    J.span = b->span;
    J.slot_idx = b->slot_idx;
    J.tris = b->tris;
    J.verts = (const uint8_t*)vertices;
    J.leaf_box = leaf_box;
    J.stride = b->stride;
    J.n = b->tri_count;
    J.chunk_offset = 0;
    J.pad = 0;
    return J;
}

luzrt_status launch_refit_batched(luzrt_ctx* ctx, const RefitJob* dev_jobs, uint32_t njobs,
                                  uint32_t total_chunks) {
    uint32_t* error = build_error_flag(ctx);
    if (!error) return fail(ctx, LUZRT_E_OUT_OF_MEMORY, "no device memory for the error flag");
    if (njobs == 0 || total_chunks == 0) return LUZRT_OK;
    RefitJob none{};
    k_refit_chunks<true><<<total_chunks, kRefitChunk, 0, ctx->stream>>>(dev_jobs, njobs, none, error);
    k_refit_spans<<<njobs, kSpanThreads, 0, ctx->stream>>>(dev_jobs, none, error);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

// world boxes, two cluster-box planes, two reference planes, slot boxes
size_t tlas_scratch_bytes(uint32_t capacity) {
    return (size_t)capacity * (4 * 6 * sizeof(float) + 2 * sizeof(uint32_t));
}

size_t tlas_lds_bytes(uint32_t m) {
    const size_t sort = (size_t)(m > 1024 ? m : 1024) * sizeof(uint64_t);
    const size_t ploc = (size_t)kPlocTile * (sizeof(float4) + sizeof(float2)) + ((size_t)m + kTlasThreads) * sizeof(uint32_t);
    return sort > ploc ? sort : ploc;
}

TlasJob tlas_job(luzrt_tlas* t, const InstIn* dev_in, uint32_t n) {
    uint32_t m = 1;
    while (m < n) m <<= 1;
    TlasJob J;
    J.in = dev_in;
    J.n = n;
    J.m = m;
    const size_t cap = t->capacity;
    J.wbox = (float*)t->scratch;
    J.cbox[0] = J.wbox + cap * 6;
    J.cbox[1] = J.cbox[0] + cap * 6;
    J.cref[0] = (uint32_t*)(J.cbox[1] + cap * 6);
    J.cref[1] = J.cref[0] + cap;
    J.nodes = t->nodes;
    J.recs = t->recs;
    J.recs_w = t->wide ? t->recs_w : nullptr;
    J.aux = t->aux;
    J.w2o_out = t->w2o;
    J.order_out = t->order;
    J.slot_box = reinterpret_cast<float*>(J.cref[1] + cap);
    return J;
}

const float* tlas_slot_boxes(const luzrt_tlas* t) {
    const size_t cap = t->capacity;
    return reinterpret_cast<const float*>(reinterpret_cast<const uint32_t*>(
                                              reinterpret_cast<const float*>(t->scratch) + cap * 18) +
                                          2 * cap);
}

luzrt_status tlas_build_device(luzrt_ctx* ctx, luzrt_tlas* t, const InstIn* dev_in, uint32_t n,
                               hipStream_t st) {
    uint32_t* error = build_error_flag(ctx);
    if (!error) return fail(ctx, LUZRT_E_OUT_OF_MEMORY, "no device memory for the error flag");
    TlasJob J = tlas_job(t, dev_in, n);
    const size_t lds = tlas_lds_bytes(J.m);
    k_tlas_build<<<1, tlas_threads(J.m), lds, st>>>(J, error);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

}  // namespace luz
