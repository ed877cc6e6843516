// tv3_internal.hpp -- internal types of the MI355X TVAL3 library.
//
// One process, one GPU, one context per host thread.  All device memory is
// owned by RAII buffers; errors travel as tv3_error exceptions inside the
// library and become tv3_status codes at the C ABI (tv3_abi.cpp).
#pragma once

#include <atomic>

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tval3_mi355x.h"

namespace tv3 {

struct tv3_error : std::runtime_error {
    tv3_status status;
    tv3_error(tv3_status s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};

#define TV3_HIP(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::tv3::tv3_error(TV3_E_HIP, std::string(#expr) + ": " +                 \
                                                  hipGetErrorString(e_));                 \
    } while (0)

#define TV3_REQUIRE(cond, st, msg)                                                        \
    do {                                                                                  \
        if (!(cond)) throw ::tv3::tv3_error(st, msg);                                     \
    } while (0)

// Device buffer of T elements (move-only).
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw tv3_error(e == hipErrorOutOfMemory ? TV3_E_NOMEM : TV3_E_HIP,
                            std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                                " bytes): " + hipGetErrorString(e));
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Pinned host buffer (hipHostMalloc).
template <typename T>
struct PinnedBuf {
    T *p = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    void alloc(size_t count) {
        if (count <= n) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        TV3_HIP(hipHostMalloc(&p, count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
};

inline size_t dtype_size(tv3_dtype t) {
    switch (t) {
    case TV3_F32: return 4;
    case TV3_F16: return 2;
    case TV3_BF16: return 2;
    case TV3_F64: return 8;
    }
    return 0;
}

enum class OpKind { CSR, DENSE, DYNAMIC, HYBRID };

// Row block of the row-blocked product kernels, read with one 32-byte
// scalar load: rows [r0, r0 + nr), entries [p0, p0 + len) of the CSR arrays
// (q0: the block's first entry in a padded entry array, = p0 when none), lg =
// log2 of the lanes per row in phase 2 (the largest power of two <= 64 with
// (1 << lg) * nr <= 256).
struct BlkDesc {
    int32_t r0, nr, p0, len, q0, lg, pad0, pad1;
};

// Descriptors of the row blocks whose first rows are blk[0 .. nb) (blk[nb] =
// end), plus one empty sentinel so that block b + 1 can always be read.
std::vector<BlkDesc> make_blk_desc(const std::vector<int32_t> &blk, const int32_t *ptr,
                                   const std::vector<int32_t> *pad_start);

// Compressed sparse rows of a matrix (also used for the CSC = CSR of A^T).
struct SparseRows {
    int64_t rows = 0, cols = 0, nnz = 0;
    tv3_dtype vtype = TV3_F32;
    DevBuf<int32_t> ptr, ind;
    DevBuf<uint8_t> val;   // nnz * dtype_size(vtype)
    int lanes = 0;         // row-group kernel: lanes per row (power of two <= 64)
    // row-blocked ("stream") kernel: blocks of consecutive rows holding at most
    // blk_nnz entries (or one longer row); blk[b] = first row of block b
    int blk_nnz = 0;       // 0: use the row-group kernel
    int64_t nblk = 0;
    DevBuf<BlkDesc> desc;  // nblk + 1 (sentinel)
};

// Partition rows into blocks for the stream kernel (host greedy over rowptr);
// starts (sorted, optional): rows where a block must begin.
void build_row_blocks(tv3_ctx *ctx, SparseRows &S, int blk_nnz,
                      const std::vector<int64_t> *starts = nullptr);

struct DynGeom {
    int32_t E = 0, R = 0, mps = 0, rv_x = 0, rv_y = 0, rv_z = 0;
    float scale = 1.f;
    tv3_weight_mode mode = TV3_WEIGHT_SIDDON;
    int64_t num_paths = 0;
    int32_t max_walk = 0;      // longest Bresenham walk (voxels) over all rays
    DevBuf<int32_t> path;      // 2 * num_paths: (emitter, receiver) global indices
    DevBuf<int32_t> ecoord;    // 3 * mps*E  (x, y, z) Bresenham
    DevBuf<int32_t> rcoord;    // 3 * mps*R
    DevBuf<float> epos, rpos;  // 3 * mps*E / 3 * mps*R world coordinates, Siddon
    float origin[3] = {0, 0, 0}, voxel[3] = {1, 1, 1};
    // host copies used to order rays for cache locality
    std::vector<float> epos_h, rpos_h;   // 3 per transducer (voxel coords for Bresenham)
    std::vector<int32_t> path_h;         // 2 per ray
};

// Spatial order of rays (internal -> natural index) for cache locality:
// transducers sorted by polar band and serpentine azimuth, rays sorted by
// (measurement position, emitter rank, receiver rank).
std::vector<int32_t> geometry_ray_order(const DynGeom &g);

// Internal index of voxel (y, x, z) in the fused solver's numbering.
//   B = 1: natural order (y fastest, then x, then z).
//   B = 4: 4x4x4 bricks of 64 voxels (y fastest inside), bricks ordered y, x, z.
//   B = 8: 8x8x8 tiles of 512 voxels ordered y, x, z; inside a tile its eight
//          4x4x4 bricks in (y, x, z) bit order, inside a brick y fastest.
// nby, nbx = bricks along y, x (B = 4); tiles along y, x (B = 8).
__host__ __device__ __forceinline__ int64_t brick_index(int B, int64_t pq, int p, int nby, int nbx,
                                                        int y, int x, int z) {
    if (B == 1) return (int64_t)z * pq + (int64_t)x * p + y;
    const int64_t in_brick = ((z & 3) << 4) + ((x & 3) << 2) + (y & 3);
    if (B == 4)
        return ((((int64_t)(z >> 2) * nbx + (x >> 2)) * nby + (y >> 2)) << 6) + in_brick;
    const int64_t tile = ((int64_t)(z >> 3) * nbx + (x >> 3)) * nby + (y >> 3);
    const int sub = (((z >> 2) & 1) << 2) | (((x >> 2) & 1) << 1) | ((y >> 2) & 1);
    return (tile << 9) + (sub << 6) + in_brick;
}

// A sparse matrix M stored for the LDS-staged panel product y = M x.  Rows
// are grouped in panels (consecutive row ranges); a panel reads x only at
// its sources -- a sorted list of indices, or a contiguous index range --
// which the product stages in LDS once; every entry addresses its source by
// a 16-bit position in that list.  Each row's entries are padded with zero
// weights to a multiple of 4 so that one 16-byte load brings a group of 4
// entries of one row; 16-bit weights are packed with their index into one
// 32-bit word (index low, weight bits high), F32 weights live in a separate
// array.  Row blocks (at most blk_nnz padded entries and 256 rows, or one
// longer row) never cross a panel.
struct Panels {
    int64_t rows = 0, npan = 0, nsrc = 0;
    int max_src = 0;           // largest source list of a panel (LDS floats)
    int blk_nnz = 0;
    int group = 4;             // rows padded to groups of 1, 2 or 4 entries
    int fmt = 0;               // 0, or 12: 16-bit weights with 12-bit indices packed,
                               // 8 entries in pk (3 words) and pw (4 words)
    int64_t stored = 4;        // stored entries (at least 4, a multiple of 4; of 8 for fmt 12)
    bool contiguous = false;   // sources of panel p: x[src_off[p] .. src_off[p + 1])
    tv3_dtype vtype = TV3_F16;
    int64_t nitem = 0;         // work items (a panel's row blocks, in chunks)
    DevBuf<int32_t> order;     // work items, decreasing work
    DevBuf<int32_t> chunk;     // per work item: panel, first block, end block
    DevBuf<int32_t> src_off;   // npan + 1 offsets into src (or index ranges)
    DevBuf<int32_t> src;       // nsrc source indices; when contiguous, empty, or the map
                               // of index ranges in super-tile order: tile t' of the
                               // range space is tile src[t'] of x (512 indices a tile)
    DevBuf<int32_t> pptr;      // rows + 1 padded row starts
    DevBuf<BlkDesc> desc;      // row blocks (+1 sentinel); p0, len in padded entries
    DevBuf<uint32_t> pk;       // padded packed entries (16-bit weights)
    DevBuf<uint16_t> ps;       // padded local indices (F32 weights)
    DevBuf<float> pv;          // padded weights (F32 weights)
    DevBuf<uint32_t> pw;       // fmt 12: the FP16 weights, 8 per 4 words
};

// Second pass of the panel form of A x: y_i = sum of the partial sums of the
// ray segments of ray i (one segment per ray and tile it crosses).
struct SegSum {
    int64_t rows = 0, nseg = 0;
    int lanes = 8;
    DevBuf<int32_t> ptr, seg;  // rows + 1, nseg: segments of each ray
    DevBuf<float> part;        // nseg partial sums (written by the first pass)
};

// Matrix-free Siddon products in tiles (dyntile.hip): the volume is cut into
// boxes (8^3 tiles, or their 4^3 bricks where a tile holds too many
// entries); a box's segments are the rays crossing it (ascending internal
// ray order).  A x walks each segment inside its box against the box's
// voxels staged in LDS, then sums each ray's segments; A^T r walks each
// segment into an LDS array of the box's entries (ray-major) and sums every
// voxel's entries through stored positions in ray order -- no atomics, the
// same bits on every run.  Only the pattern's positions are stored (16 bits
// per entry); every weight is recomputed by the traversal of siddon.hpp.
struct DynPan {
    int32_t v0;          // first internal voxel of the box (of its first tile)
    int32_t bx, by, bz;  // box corner in voxel coordinates (x, y, z)
    int32_t bs;          // box edge along x and z: 16 (2 x 2 x 2 tiles), 8 (tiles) or 4
                         // (a brick)
    int32_t ey;          // box edge along y: bs, or 16 for two tiles consecutive in y
    int32_t s0, s1;      // its segments
    int32_t ne;          // entries of its segments
};

// One direction's boxes and segments.
struct DynBoxes {
    int64_t npan = 0, nseg = 0;      // work items (pan), segments
    int32_t emax = 0, smax = 0;      // most entries / segments of one box
    DevBuf<DynPan> pan;              // A^T r: boxes; A x: work items (a box, part of its
                                     // segments)
    DevBuf<int32_t> order;           // work items, decreasing work
    DevBuf<int32_t> seg_ray;         // internal ray of each segment
    DevBuf<int32_t> seg_eoff;        // first entry of each segment in its box (A^T)
    DevBuf<uint16_t> seg_c0;         // box-local first voxel of each segment: x | y << 4 | z << 8
};

struct DynTiles {
    int64_t nnz = 0;
    int32_t emax_cap = 0;                 // A^T box limit: entries in LDS
    int32_t threads_T = 512;              // workgroup of the A^T product
    float org[3] = {0, 0, 0}, h[3] = {1, 1, 1};
    int32_t N[3] = {0, 0, 0};        // voxels along x, y, z
    int32_t tile_dx = 0, tile_dz = 0;  // internal offset of the next 8^3 tile along x, z
    DynBoxes fwd;                    // A x: 2 x 2 x 2 tiles, tile pairs along y, tiles or
                                     // bricks (B = 4)
    DynBoxes bwd;                    // A^T r: tiles, split into bricks where a tile's
                                     // entries exceed the LDS limit
    DevBuf<float> ray;               // 12 per internal ray: sid_setup's A[3], L, B[3],
                                     // amin, amax (3 unused)
    DevBuf<uint32_t> pptr;           // n + 1: padded starts of the voxel rows
    DevBuf<uint16_t> pos;            // per voxel row its entries' positions in its bwd
                                     // box, ray order, padded to 4 with the position ne
    DevBuf<float> xin, yout;         // natural <-> internal order staging (max(n, m))
};

// The fused solver's internal layout of a sparse operator: voxels numbered by
// brick_index (B = 8 when every dimension is a multiple of 8, else 4 when a
// multiple of 4, else 1), rays in a locality order; A and its CSC in that
// numbering, and their panel forms where those were tuned in.
// What a fused solve keeps for the next solve with the same layout: state
// buffers, control block, partial sums and its instantiated graphs.
struct FusedSession {
    virtual ~FusedSession() = default;
};

struct FusedLayout {
    int32_t dims[3] = {0, 0, 0};
    int B = 1, nby = 0, nbx = 0;
    SparseRows A, AT;
    // panel forms: A^T x directly (rows = voxels, sources = rays); A x as
    // per-tile segment sums (rows = segments, sources = the tile's voxels)
    // followed by the per-ray sum RS
    Panels PT, PA;
    SegSum RS;
    bool panels_T = false, panels_A = false;
    int unit_A = 512;                 // voxels per panel of A's panel form (1, 2, 4 or 8 tiles)
    int unit_T = 512;                 // voxels per panel of A^T's panel form (1 or 2 tiles)
    int64_t npan = 0, nseg = 0;
    int max_seg = 0;
    DevBuf<int32_t> v_nat2int, v_int2nat;   // voxel maps (n)
    DevBuf<int32_t> r_int2nat;              // ray map (m): internal ray -> natural ray
    // the solver's buffers and graphs for this layout (dropped with it, and
    // whenever its launch shapes change)
    std::unique_ptr<FusedSession> session;
    // a matrix-free Siddon operator: its tile products (RS is their second
    // pass); A and AT then keep only their sizes
    std::unique_ptr<DynTiles> dyn;
};

}  // namespace tv3

namespace tv3 {
// State a context keeps between calls (the last batched solve's layouts,
// tuned shapes and graphs).
struct CtxSession {
    virtual ~CtxSession() = default;
};
// serial number of an operator, never reused: a cached session recognises
// its operators by it, not by their addresses
inline uint64_t next_op_uid() {
    static std::atomic<uint64_t> counter{0};
    return ++counter;
}
}  // namespace tv3

struct tv3_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cus = 256;
    size_t lds_per_block = 65536;   // largest LDS allocation of one workgroup
    // reduction scratch
    tv3::DevBuf<double> red_partials;
    tv3::PinnedBuf<double> red_host;
    tv3::DevBuf<uint8_t> scratch;   // general workspace (grown on demand)
    std::unique_ptr<tv3::CtxSession> batch_session;   // the last batched solve, for reuse
};

struct tv3_op {
    uint64_t uid = tv3::next_op_uid();
    tv3_ctx *ctx = nullptr;
    int device = 0;                  // device of ctx (destruction does not read ctx)
    tv3::OpKind kind = tv3::OpKind::CSR;
    int64_t m = 0, n = 0;
    // CSR / HYBRID: A rows and A^T rows (the CSC)
    tv3::SparseRows A, AT;
    // DENSE
    tv3_dtype dtype = TV3_F32;
    tv3::DevBuf<uint8_t> dense;
    int64_t lda = 0;
    tv3::DevBuf<uint8_t> dense_oN, dense_oT;   // A and A^T in MFMA operand order (multi-RHS
                                               // products), built on demand
    // DYNAMIC / HYBRID
    tv3::DynGeom geo;
    // CSR built from a geometry: the locality order of its rays (may be empty)
    std::vector<int32_t> ray_order;
    // fused solver layout, built on first use for a given dims
    std::unique_ptr<tv3::FusedLayout> fused;
};

namespace tv3 {

// ---------------------------------------------------------------- launchers
// sparse (spmv.hip)
// acc64: FP64 accumulation of each output entry, rounded once (FP32 vectors)
void spmv_csr(const SparseRows &S, const float *x, float *y, hipStream_t s, int lanes = 0,
              bool acc64 = false);
void spmv_csr(const SparseRows &S, const double *x, double *y, hipStream_t s, int lanes = 0,
              bool acc64 = true);
int default_lanes(int64_t rows, int64_t nnz);
// Stable device transpose of a CSR (rows ordered ascending within each column).
void build_transpose(tv3_ctx *ctx, const SparseRows &A, SparseRows &AT);
// exclusive scan of cnt[0..n) into out[0..n] (out[n] = total); returns total
int64_t exclusive_scan_i32(tv3_ctx *ctx, const int32_t *cnt, int32_t *out, int64_t n);

// dense (dense.hip)
void dense_gemv(tv3_op *op, bool trans, const float *x, float *y, hipStream_t s,
                bool acc64 = false);
void dense_gemv(tv3_op *op, bool trans, const double *x, double *y, hipStream_t s,
                bool acc64 = true);
// y = A x as dense_gemv (FP32 accumulation) with the split-K partials in the
// caller's workspace ws of ws_bytes; a larger need is rejected
void dense_gemv_ws(tv3_op *op, const float *x, float *y, hipStream_t s, void *ws,
                   size_t ws_bytes);
// the allocation-free path (noalloc.hip): per-device address of its scratch
// array, looked up once per device with that device current
void noalloc_setup(int device);

// dense multi-right-hand-side products on the matrix cores (dense_mfma.hip):
// Y = M X (- sub), M BF16 R x K in MFMA operand order (dense_operand_order),
// X/Y column-major, G <= 64; a
// non-null gate in device memory turns the launch into a no-op while *gate
// is 0 (graph-captured launches that only some rounds need)
int gemm_rhs_rows();   // rows per workgroup (R must be a multiple)
bool gemm_rhs_shape_ok(int64_t R, int64_t K, int splits);
void gemm_rhs(const uint16_t *M, int64_t R, int64_t K, const float *X, int64_t ldx,
              int G, float *Y, int64_t ldy, const float *sub, int splits, float *part,
              hipStream_t s, const int32_t *gate = nullptr);
size_t gemm_rhs_scratch(int64_t R, int splits);   // floats
void dense_operand_order(tv3_op *op, hipStream_t s);   // builds op->dense_oN / _oT once

// dynamic (dynamic.hip)
void dynamic_apply(tv3_op *op, bool trans, const float *x, float *y, hipStream_t s,
                   bool acc64 = false);
void dynamic_apply(tv3_op *op, bool trans, const double *x, double *y, hipStream_t s,
                   bool acc64 = true);
void dynamic_materialize(tv3_op *op, tv3_dtype vt, SparseRows &A);

// generic apply: y = op x (trans = 0) or op^T x
void op_apply(tv3_op *op, bool trans, const float *x, float *y, hipStream_t s,
              bool acc64 = false);
void op_apply(tv3_op *op, bool trans, const double *x, double *y, hipStream_t s,
              bool acc64 = true);

// internal layout (layout.hip)
// the fused layout of op for a volume shape, built on first use (launch
// shapes timed on this GPU unless tune is false: then the row kernels'
// defaults, no panel forms)
FusedLayout &fused_layout(tv3_op *op, const int32_t dims[3], bool tune = true);
// the layout's numbering only: voxel and ray maps, A and its CSC renumbered
// (no panel forms, no tuned launch shapes)
void renumber_operator(tv3_op *op, const int32_t dims[3], FusedLayout *L);
// the same for the rows A (m rays, natural order) of any operator
void renumber_rows(tv3_ctx *ctx, const SparseRows &A, const std::vector<int32_t> &ray_order,
                   const int32_t dims[3], FusedLayout *L);
// tile products of a matrix-free Siddon operator (dyntile.hip): the layout
// for dims = (rv_y, rv_x, rv_z) with every dimension a multiple of 4, built
// at creation; null when the geometry does not allow it
std::unique_ptr<FusedLayout> build_dyn_layout(tv3_op *op);
// y = A x / A^T x of a matrix-free operator by its tile products, natural
// order in and out; false when it has none (FP32 vectors)
bool dyn_tile_apply(tv3_op *op, bool trans, const float *x, float *y, hipStream_t s);
// panel forms of a prebuilt A^T (rows = voxels in tiles of 512 rows, sources =
// rays), one per row-block size in blks; false when a panel would need more
// distinct rays than its local indices address or the weights are not F32,
// F16 or BF16.  16-bit weights in blocks of 2048 or 4096 are packed (12-bit
// indices over panels of at most 4096 rays)
bool build_transposed_panels(tv3_ctx *ctx, const SparseRows &AT, const std::vector<int> &blks,
                             std::vector<Panels> &out);
// A volume of 8^3 tiles inside the rows of a prebuilt A^T: its first row and
// tile counts along y, x, z
struct TileVolume {
    int64_t row0;
    int nby, nbx, nbz;
};
// the two-pass panel form of A from a prebuilt A^T (rows of A^T = voxels in
// units of 512), one per row-block size; seg_starts[k] = first segment of
// the voxel range starting at row_starts[k] (row_starts ascending, multiples
// of 512; seg_starts gets one more entry: the segment count).  With boxes,
// panels over 16^3 voxels of those volumes (each starting at a multiple of
// 4096 rows with even tile counts; false otherwise), packed for 16-bit
// weights in blocks of 2048 or 4096
bool build_forward_panels(tv3_ctx *ctx, const SparseRows &AT, const std::vector<int> &blks,
                          const std::vector<int64_t> &row_starts, std::vector<Panels> &P,
                          std::vector<SegSum> &R, std::vector<int32_t> &seg_starts,
                          const std::vector<TileVolume> *boxes = nullptr);
// y = A x or A^T x with the layout's kernels, natural order in and out
void layout_matvec(tv3_op *op, const int32_t dims[3], bool trans, const float *x, float *y,
                   hipStream_t s);
// override the layout's launch shapes (tv3_op_layout_force)
void layout_force(tv3_op *op, const int32_t dims[3], int code_N, int code_T, int pan_nnz,
                  int panels);
// HBM bandwidth probes (GB/s): streaming copy (bytes read + bytes written)
// and streaming read of a buffer of the given size, median of reps launches
void bandwidth_probe(tv3_ctx *ctx, int64_t bytes, int reps, double *copy_gbps, double *read_gbps);
// average duration (us) of one layout product kernel launch
double layout_time(tv3_op *op, const int32_t dims[3], bool trans, int reps);
void permute_gather(const float *in, const int32_t *map, float *out, int64_t n, hipStream_t s);
void permute_scatter(const float *in, const int32_t *map, float *out, int64_t n, hipStream_t s);

// solvers
template <typename T>
void solve_plain(tv3_ctx *ctx, tv3_op *op, const tv3_opts &o, const int32_t dims[3],
                 const T *b_dev, T *U_dev, const T *Ut_dev, tv3_info *info, tv3_history *hist);
void solve_fused(tv3_ctx *ctx, tv3_op *op, const tv3_opts &o, const int32_t dims[3],
                 const float *b_dev, float *U_dev, const float *Ut_dev, tv3_info *info,
                 tv3_history *hist);
void solve_batched(tv3_ctx *ctx, int P, tv3_op *const *ops, const int32_t *dims3,
                   const tv3_opts &o, const float *const *b_dev, float *const *U_dev,
                   const float *const *Ut_dev, tv3_info *infos, tv3_history *hists,
                   tv3_batch_info *binfo);

// stencil building blocks (vec.hip), float and double
template <typename T>
void launch_D(const int32_t dims[3], const T *U, T *Ux, T *Uy, T *Uz, hipStream_t s);
template <typename T>
void launch_Dt(const int32_t dims[3], const T *X, const T *Y, const T *Z, T *res, hipStream_t s);
template <typename T>
double shrink_host(tv3_ctx *ctx, int64_t n, int iso, const T *Ux, const T *Uy, const T *Uz,
                   const T *sx, const T *sy, const T *sz, double beta, T *Wx, T *Wy, T *Wz,
                   hipStream_t s);

// shrinkage as a function of V = DU - sigma/beta and its VJP (autodiff.hip)
template <typename T>
void shrink_apply(int64_t n, int iso, double t, const T *Vx, const T *Vy, const T *Vz, T *Wx,
                  T *Wy, T *Wz, T *s, const uint8_t *mask_in, uint8_t *mask_out,
                  hipStream_t st);
template <typename T>
void shrink_adjoint(int64_t n, int iso, double t, const T *Vx, const T *Vy, const T *Vz,
                    const T *gWx, const T *gWy, const T *gWz, const T *gs, T *gVx, T *gVy,
                    T *gVz, hipStream_t st);

}  // namespace tv3
