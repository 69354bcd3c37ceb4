/*
 * gpugraphs.h -- C ABI of the MI355X (gfx950) semiring sparse kernels.
 *
 * Drop-in boundary for the hot path of GPUGraphs.jl: the storage types, the
 * semiring SpMV / SpMM / SpMSpV kernels and the BFS / Bellman-Ford loops built
 * on them.  Every entry point takes plain pointers and sizes; device buffers
 * are HIP device pointers; indices are 1-based Int32 exactly as in the
 * reference's structs, so a Julia wrapper passes its struct fields unchanged.
 *
 * Conventions
 *  - Every function returns a gg_status.  GG_OK = 0.  The other codes name the
 *    Julia exception the reference throws for the same condition:
 *      GG_ERR_DIMENSION  -> DimensionMismatch  (src/spmv/CSR.jl:374-379,400-402)
 *      GG_ERR_ARGUMENT   -> ArgumentError      (src/spmv/CSR.jl:381-387,
 *                                                src/storage/CSR.jl:48-63)
 *      GG_ERR_ASSERTION  -> AssertionError     (src/spmm/CSR.jl:131-133,161)
 *      GG_ERR_BOUNDS     -> BoundsError        (src/storage/CSR.jl:155-157)
 *    gg_last_error() returns the message of the last failure on this thread.
 *  - Kernels are enqueued on the context's stream and return without waiting
 *    (the reference launches asynchronously and the caller synchronises,
 *    test/spmv.jl:24-25).  Algorithm entry points (BFS, Bellman-Ford, the
 *    batched forms) return after their result is complete.
 *  - Operators are codes, one per GPUGraphs_* function of src/algebra.jl.
 */
#ifndef GPUGRAPHS_H
#define GPUGRAPHS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GG_ABI_VERSION 1

typedef enum gg_status {
  GG_OK = 0,
  GG_ERR_DIMENSION = 1,
  GG_ERR_ARGUMENT = 2,
  GG_ERR_ASSERTION = 3,
  GG_ERR_BOUNDS = 4,
  GG_ERR_UNSUPPORTED = 5,  /* operator / type combination without a kernel */
  GG_ERR_HIP = 6,          /* a HIP runtime call failed */
  GG_ERR_NOMEM = 7,
  GG_ERR_NOT_CONVERGED = 8,/* iteration cap reached (negative cycle) */
  GG_ERR_BUDGET = 9        /* request exceeds the serving memory budget */
} gg_status;

/* Element types.  GG_BOOL is one byte holding 0 or 1 (Julia Bool). */
typedef enum gg_dtype {
  GG_BOOL = 1,
  GG_INT32 = 2,
  GG_INT64 = 3,
  GG_FLOAT32 = 4,
  GG_FLOAT64 = 5,
  GG_UINT64 = 6
} gg_dtype;

/* Operators of src/algebra.jl, one code per function. */
typedef enum gg_op {
  GG_OP_FIRST = 1,     /* GPUGraphs_first     algebra.jl:31 */
  GG_OP_SECOND = 2,    /* GPUGraphs_second    algebra.jl:32 */
  GG_OP_ANY = 3,       /* GPUGraphs_any       algebra.jl:35 (max) */
  GG_OP_PAIR = 4,      /* GPUGraphs_pair      algebra.jl:40 */
  GG_OP_ADD = 5,       /* GPUGraphs_add       algebra.jl:43 */
  GG_OP_SAFE_ADD = 6,  /* GPUGraphs_safe_add  algebra.jl:47 */
  GG_OP_MUL = 7,       /* GPUGraphs_mul       algebra.jl:55 */
  GG_OP_MIN = 8,       /* GPUGraphs_min       algebra.jl:65 */
  GG_OP_MAX = 9,       /* GPUGraphs_max       algebra.jl:68 */
  GG_OP_LOR = 10,      /* GPUGraphs_lor       algebra.jl:77 */
  GG_OP_LAND = 11,     /* GPUGraphs_land      algebra.jl:80 */
  GG_OP_BOR = 12,      /* GPUGraphs_bor       algebra.jl:95 */
  GG_OP_BAND = 13,     /* GPUGraphs_band      algebra.jl:98 */
  GG_OP_SECONDI = 14   /* GPUGraphs_secondi   algebra.jl:109 (B_i * y) */
} gg_op;

typedef enum gg_format {
  GG_CSR = 1,   /* SparseGPUMatrixCSR  src/storage/CSR.jl:14-24 */
  GG_SELL = 2,  /* SparseGPUMatrixSELL src/storage/SELL.jl:3-18 */
  GG_CSC = 3    /* SparseGPUMatrixCSC  src/storage/CSC.jl:2-12 */
} gg_format;

typedef enum gg_layout {
  GG_COL_MAJOR = 0, /* Julia Matrix layout, the reference's SpMM operands */
  GG_ROW_MAJOR = 1  /* one contiguous row of k values per vertex */
} gg_layout;

/*
 * Non-owning view of a sparse matrix whose arrays live in device memory.
 * CSR:  ptr = rowptr[m+1], idx = colval[nnz], val = nzval[nnz]
 * CSC:  ptr = colptr[n+1], idx = rowval[nnz], val = nzval[nnz]
 * SELL: ptr = slice_ptr[nslices+1], idx = colval[n_stored] (pad -1),
 *       val = nzval[n_stored] (pad 0).  Entry k of slot r of slice s sits at
 *       slice_ptr[s] + k*slice_size + r (1-based), src/storage/SELL.jl:128-161.
 *       perm == NULL and n_heavy == 0 is exactly the reference layout.
 *       perm[r] (1-based, 0 = empty slot) is the matrix row held by slot r
 *       when rows were sorted by length inside windows of sigma rows;
 *       slot_len[r] is that slot's entry count.  Rows longer than the SELL
 *       bound are kept out of the slices in a CSR "heavy" partition that the
 *       same launch processes in segments of heavy_seg_len entries.
 */
typedef struct gg_matrix {
  int32_t format;     /* gg_format */
  int32_t val_dtype;  /* gg_dtype of val / heavy_val */
  int64_t m, n, nnz;
  const int32_t* ptr;
  const int32_t* idx;
  const void* val;
  /* SELL-C-sigma only */
  int64_t slice_size;
  int64_t nslices;
  int64_t n_stored;
  const int32_t* perm;      /* [nslices*slice_size] or NULL */
  const int32_t* slot_len;  /* [nslices*slice_size] or NULL */
  int64_t n_heavy;
  const int32_t* heavy_row; /* [n_heavy] 1-based rows */
  const int64_t* heavy_ptr; /* [n_heavy+1] 0-based offsets */
  const int32_t* heavy_seg; /* [n_heavy+1] prefix count of segments */
  int64_t heavy_seg_len;
  int64_t heavy_nseg;       /* heavy_seg[n_heavy] */
  const int32_t* heavy_col; /* 1-based columns */
  const void* heavy_val;
  int64_t flags;            /* gg_matrix_flags */
  /* GG_FLAG_INTERNAL_IDS: columns (colval, heavy_col) are internal ids
   * 1..nslices*slice_size + n_heavy (slot index, then n_slots + heavy index);
   * vinv[original row] is its internal id. */
  const int32_t* vinv;
} gg_matrix;

typedef enum gg_matrix_flags {
  /* Every stored value is Bool true or an odd integer, so (value & x) == x
   * for x in {0,1}: or-and BFS may skip loading values. */
  GG_FLAG_UNIT_PATTERN = 1,
  /* A relabelled SELL (gg_sell_build with GG_SELL_RELABEL). */
  GG_FLAG_INTERNAL_IDS = 2,
  /* The pattern is symmetric: row v of A_T also lists the out-neighbours of
   * v, so traversals may expand a small frontier top-down (push) over the
   * same arrays; used with GG_FLAG_UNIT_PATTERN by multi-source BFS. */
  GG_FLAG_SYMMETRIC = 4
} gg_matrix_flags;

/* Host-side SELL-C-sigma arrays produced by gg_sell_build (owned by the
 * library until gg_sell_host_free). */
typedef struct gg_sell_host {
  int64_t m, n, nnz, slice_size, sigma, nslices, n_stored, n_slots;
  int32_t val_dtype;
  int32_t* slice_ptr;  /* [nslices+1] */
  int32_t* colval;     /* [n_stored] */
  void* nzval;         /* [n_stored] */
  int32_t* perm;       /* [n_slots] or NULL (identity) */
  int32_t* slot_len;   /* [n_slots] or NULL */
  int64_t n_heavy, heavy_nnz, heavy_seg_len, n_seg;
  int32_t* heavy_row;  /* [n_heavy] */
  int64_t* heavy_ptr;  /* [n_heavy+1] */
  int32_t* heavy_seg;  /* [n_heavy+1] */
  int32_t* heavy_col;  /* [heavy_nnz] */
  void* heavy_val;     /* [heavy_nnz] */
  int32_t* vinv;       /* [m] original row -> internal id, or NULL (GG_SELL_RELABEL) */
  int64_t n_internal;  /* size of the internal id space */
} gg_sell_host;

/* gg_sell_build flags */
typedef enum gg_sell_flags {
  /* Relabel vertices by decreasing degree: rows are laid out in that order
   * (the sigma windows become moot), a light row's internal id is its slot,
   * a heavy row's is n_slots + its heavy index, and stored columns use
   * internal ids.  Entry points gather x into internal order first; results
   * are written at the original row ids.  Needs a square matrix. */
  GG_SELL_RELABEL = 1,
  /* Store each row's entries by increasing stored column (with relabelling:
   * light rows by decreasing degree, heavy rows last) instead of the input
   * order.  Folds then visit a row's terms in that order: same results for
   * min / max / or / and, a reordered sum for plus on floats. */
  GG_SELL_SORT_ENTRIES = 2,
  /* With GG_SELL_RELABEL: rows of equal length are laid out in reverse
   * Cuthill-McKee order (gg_rcm_order) instead of row order, so neighbouring
   * light rows share more of their gathered columns.  Host builder only
   * (gg_sell_build); gg_sell_plan_device rejects it. */
  GG_SELL_RCM = 4
} gg_sell_flags;

typedef struct gg_ctx gg_ctx;

/* The SELL builder on the device (no counterpart in the reference, which
 * builds on the host, src/storage/SELL.jl:99-181): the layouts of
 * gg_sell_build, array for array, from a CSR already in device memory.
 * gg_sell_plan_device orders the rows and returns every output size; the
 * caller allocates device arrays of those sizes in a gg_matrix -- ptr =
 * slice_ptr[nslices+1], idx = colval[n_stored], val = nzval[n_stored];
 * perm / slot_len [n_slots] when has_perm; heavy_row [n_heavy], heavy_ptr
 * [n_heavy+1] (Int64), heavy_seg [n_heavy+1], heavy_col / heavy_val
 * [heavy_nnz]; vinv [m] when has_vinv -- and gg_sell_fill_device writes them
 * and the matrix's scalar fields.  The CSR arrays must stay alive until the
 * fill; the plan is freed with gg_sell_plan_free. */
typedef struct gg_sell_plan gg_sell_plan;
typedef struct gg_sell_sizes {
  int64_t slice_size, nslices, n_stored, n_slots;
  int64_t n_heavy, heavy_nnz, n_seg, n_internal;
  int32_t has_perm, has_vinv;
} gg_sell_sizes;

/* ---------------------------------------------------------------- runtime */
int gg_abi_version(void);
const char* gg_status_string(int status);
const char* gg_last_error(void);
/* stream: a hipStream_t, or NULL for the device's null stream. */
int gg_ctx_create(int device, void* stream, gg_ctx** out);
int gg_ctx_destroy(gg_ctx* ctx);
int gg_ctx_synchronize(gg_ctx* ctx);
int gg_ctx_set_stream(gg_ctx* ctx, void* stream);
/* Device memory for callers without an allocator of their own (Julia's
 * KernelAbstractions.allocate counterpart). */
int gg_malloc(gg_ctx* ctx, size_t bytes, void** out);
int gg_free(gg_ctx* ctx, void* ptr);
int gg_memcpy_htod(gg_ctx* ctx, void* dst, const void* src, size_t bytes);
int gg_memcpy_dtoh(gg_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Timing with HIP events on the context stream, when enabled (it adds a
 * synchronisation after each timed launch):
 *  - gg_ctx_last_kernel_ms: milliseconds of the most recent entry point;
 *  - gg_ctx_hot_kernel_stats: summed milliseconds and count of the hot-kernel
 *    launches (the SpMV / SpMM / traversal kernel of each entry point, not
 *    its fills and bookkeeping) since the last reset. */
int gg_ctx_enable_timing(gg_ctx* ctx, int on);
int gg_ctx_last_kernel_ms(gg_ctx* ctx, double* ms);
int gg_ctx_hot_kernel_stats(gg_ctx* ctx, double* total_ms, int64_t* launches, int reset);
/* Level / iteration loops (gg_bfs_distances, gg_msbfs_distances,
 * gg_shortest_path[_multi]) run as HIP graphs of `steps` captured steps
 * (rounded up to even), relaunched until a device-side convergence flag is
 * set: one host round trip per `steps` levels instead of one per level.  The
 * graphs are cached in the context per matrix and workspace.  0 (the
 * default) launches every step from the host.  The reference's loops sync
 * on the host every level (bfs.jl:60) or every 8 iterations
 * (shortest_path.jl:71-76). */
int gg_ctx_set_graph_steps(gg_ctx* ctx, int steps);
/* Lanes per row of the CSR SpMV, BFS and Bellman-Ford kernels (a power of
 * two in 1:64; 0 = the next power of two >= the average row length).  The
 * reference runs one work-item per row (src/spmv/CSR.jl:31-52). */
int gg_ctx_set_csr_lanes(gg_ctx* ctx, int lanes);
/* Direction switch on symmetric patterns (GG_FLAG_SYMMETRIC): a level of
 * gg_msbfs_distances (unit pattern) runs top-down when its frontier's
 * out-edges times alpha are below the matrix entries, and gg_bfs_packed
 * expands its first levels top-down while that holds (default 16; 0 =
 * never). */
int gg_ctx_set_push_threshold(gg_ctx* ctx, double alpha);
/* Sparse iterations of gg_shortest_path_multi on a symmetric relabelled
 * SELL: an iteration relaxes only the rows next to the rows that fell in the
 * previous one when those rows' out-edges times alpha are below the matrix
 * entries (default 16; 0 = never). */
int gg_ctx_set_sssp_threshold(gg_ctx* ctx, double alpha);

/* Host-side consistency check of a descriptor (no device access): the checks
 * every entry point runs before launching. */
int gg_matrix_check(const gg_matrix* A);
int gg_sell_plan_device(gg_ctx* ctx, int64_t m, int64_t n, const int32_t* rowptr,
                        const int32_t* colval, const void* nzval, int val_dtype,
                        int64_t slice_size, int64_t sigma, int64_t heavy_threshold,
                        int64_t heavy_seg_len, int flags, gg_sell_plan** plan,
                        gg_sell_sizes* sizes);
int gg_sell_fill_device(gg_sell_plan* plan, gg_matrix* out);
void gg_sell_plan_free(gg_sell_plan* plan);
/* SELL slots the kernels walk: nslices*slice_size with a permutation, else
 * min(m, nslices*slice_size) (identity layout); 0 for CSR / CSC. */
int gg_matrix_slots(const gg_matrix* A, int64_t* n_slots);

/* ---------------------------------------------------------------- storage */
/* Validation of a CSR / CSC triple held in host memory; same checks and order
 * as the inner constructors src/storage/CSR.jl:48-63, src/storage/CSC.jl:31-46. */
int gg_csr_check(int64_t m, int64_t n, const int64_t* rowptr, int64_t rowptr_len,
                 const int64_t* colval, int64_t colval_len, int64_t nzval_len);
/* SELL inner-constructor checks, src/storage/SELL.jl:42-52. */
int gg_sell_check(int64_t n, int64_t nslices, int64_t n_stored, int64_t slice_ptr_len,
                  const int64_t* colval, int64_t colval_len, int64_t nzval_len);
/* SELL-C-sigma builder over a host CSR (1-based).  sigma <= 1 and
 * heavy_threshold <= 0 give the reference layout, src/storage/SELL.jl:99-181;
 * sigma > 1 sorts rows by length inside windows of sigma rows;
 * heavy_threshold > 0 moves rows longer than it into the heavy partition. */
/* Reverse Cuthill-McKee order of a square CSR pattern (host arrays, 1-based):
 * order[i] = the vertex at position i.  Components from their lowest-degree
 * vertex, neighbours by increasing degree, the sequence reversed.  A layout
 * aid with no counterpart in the reference. */
int gg_rcm_order(int64_t m, const int32_t* rowptr, const int32_t* colval, int32_t* order);
int gg_sell_build(int64_t m, int64_t n, const int32_t* rowptr, const int32_t* colval,
                  const void* nzval, int val_dtype, int64_t slice_size, int64_t sigma,
                  int64_t heavy_threshold, int64_t heavy_seg_len, int flags,
                  gg_sell_host** out);
void gg_sell_host_free(gg_sell_host* h);

/* ------------------------------------------------------------- generators */
/* Seeded synthetic graphs as host CSR of A_T (row v = in-neighbours of v,
 * 1-based, sorted, deduplicated).  Every random draw is a hash of the seed and
 * the ids it decides, so a graph depends only on its parameters.  The
 * reference draws its inputs with Julia's sprand / Graphs.jl generators
 * (test/spmv.jl:12, test/bfs.jl:11-12, benchmark/benchmarks.jl:64-65). */
typedef struct gg_csr_host {
  int64_t m, n, nnz;
  int32_t val_dtype;
  int32_t* rowptr;  /* [m+1] */
  int32_t* colval;  /* [nnz] */
  void* nzval;      /* [nnz] */
} gg_csr_host;
void gg_csr_host_free(gg_csr_host* h);
/* values: 0 = unit (Bool true / 1), 1 = U[0,1) */
/* R-MAT (Chakrabarti et al. 2004): 2^scale vertices, edge_factor*2^scale
 * draws with quadrant probabilities a, b, c, 1-a-b-c; ids randomly permuted
 * when permute; symmetrised; self-loops and duplicates removed. */
int gg_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                int permute, int val_dtype, int values, gg_csr_host** out);
/* rows x cols 4-neighbour grid, both directions with independent integer
 * weights in [wmin, wmax], plus round(long_range_frac * grid edges) random
 * vertex pairs (u != v) added in both directions with independent weights. */
int gg_gen_grid(int64_t rows, int64_t cols, double long_range_frac, int32_t wmin, int32_t wmax,
                uint64_t seed, int val_dtype, gg_csr_host** out);
/* n x n matrix whose rows hold exactly per_row distinct uniformly drawn
 * columns, values U[0,1). */
int gg_gen_uniform_rows(int64_t n, int64_t per_row, uint64_t seed, int val_dtype,
                        gg_csr_host** out);
/* Uniform random graph: n*degree/2 uniform vertex pairs, symmetrised.
 * values: 0 unit, 1 U[0,1), 2 integer weights in [wmin, wmax]. */
int gg_gen_random_graph(int64_t n, int64_t degree, uint64_t seed, int val_dtype, int values,
                        int32_t wmin, int32_t wmax, gg_csr_host** out);
/* MatrixMarket coordinate file (real, integer or pattern; general, symmetric
 * or skew-symmetric) -> CSR of the stored matrix, or of its transpose when
 * transpose != 0; duplicates add up (pattern: kept once); a symmetric file is
 * expanded to both triangles.  val_dtype 0 keeps the file's type (pattern ->
 * Bool, integer -> Int64, real -> Float64); *symmetric (optional) is set to 1
 * for a symmetric or skew-symmetric file.  The reference reads SuiteSparse
 * matrices from the network (benchmark/benchmarks.jl:227-231,262-263); this
 * reads the same matrices from local files. */
int gg_read_matrix_market(const char* path, int val_dtype, int transpose, gg_csr_host** out,
                          int* symmetric);
/* Edge list: one "u v" or "u v w" per line ('#' or '%' starts a comment),
 * vertices counted from `base` (0 or 1).  Output is A_T (row v lists the
 * tails u of the edges u -> v, the generators' convention); directed == 0
 * adds v -> u for every edge; a repeated edge keeps its smallest weight;
 * unweighted edges have value 1 (Bool true).  No counterpart in the
 * reference. */
int gg_read_edge_list(const char* path, int base, int directed, int val_dtype,
                      gg_csr_host** out);
/* Block-diagonal packing of n_graphs square CSRs (A_T each, one value type):
 * graph g owns rows and columns offsets[g] .. offsets[g+1]-1 (0-based;
 * offsets has n_graphs+1 entries), each block padded with isolated vertices to
 * a multiple of 256 rows.  No counterpart in the reference, which runs one
 * graph per call (bfs.jl:1-111); this is the input of gg_bfs_packed. */
int gg_pack_graphs(int n_graphs, const gg_csr_host* const* graphs, gg_csr_host** out,
                   int64_t* offsets);

/* ---------------------------------------------------------------- kernels */
/* C = accum(C, A (add.mul) B) on rows selected by mask (mask[r] != 0) and the
 * 1-based row range [range_start, range_start+range_len).  mask == NULL: all
 * rows; range_len < 0: no range.  Replaces gpu_spmv! for CSR, SELL and CSC:
 * src/spmv/CSR.jl:2-29, src/spmv/SELL.jl:2-29, src/spmv/CSC.jl:23-61. */
int gg_spmv(gg_ctx* ctx, void* C, int c_dtype, int64_t c_len, const gg_matrix* A,
            const void* B, int b_dtype, int64_t b_len, int add, int mul, int accum,
            const void* mask, int mask_dtype, int64_t mask_len, int64_t range_start,
            int64_t range_len);

/* Sparse-mask SpMV: C[r] = accum(C[r], A[r,:] (add.mul) B) for each row r of
 * rows[0..n_rows) (1-based, distinct; entries outside 1..m are skipped).
 * Replaces sparse_masked_csr_spmv_kernel!, src/spmv/CSR.jl:90-115, which no
 * dispatch of the reference reaches (gpu_spmv! takes a dense mask there).  CSR
 * only; SELL and CSC return GG_ERR_UNSUPPORTED. */
int gg_spmv_rows(gg_ctx* ctx, void* C, int c_dtype, int64_t c_len, const gg_matrix* A,
                 const void* B, int b_dtype, int64_t b_len, int add, int mul, int accum,
                 const int32_t* rows, int64_t n_rows);

/* Dense-operand SpMM, C[m x k] = accum(C, A (add.mul) B[n x k]).  Replaces
 * gpu_spmm!, src/spmm/CSR.jl:111-204 and src/spmm/SELL.jl:84-124.  The
 * reference passes Julia (column-major) matrices; GG_ROW_MAJOR is the layout
 * of the batched multi-source kernels.  A CSC A (src/spmm/CSC.jl is a
 * placeholder) runs the k-column form of the CSC SpMV: C += A * B by atomic
 * adds, add and accum not applied, one element type (Float32, Float64,
 * Int32), no mask or range. */
int gg_spmm(gg_ctx* ctx, void* C, int c_dtype, int64_t c_rows, int64_t c_cols,
            int c_layout, const gg_matrix* A, const void* B, int b_dtype, int64_t b_rows,
            int64_t b_cols, int b_layout, int add, int mul, int accum, const void* mask,
            int mask_dtype, int64_t mask_len, int64_t range_start, int64_t range_len);

/* Masked sparse-matrix x sparse-vector product (the kernel src/spmspv.jl:2-54
 * intends): for every stored b_j, every stored A[i,j] contributes
 * mul(A[i,j], b_j) to c_i under add; rows i with mask[i] == 0 (or != 0 when
 * mask_complement) are skipped.  A is CSC (columns expand directly) or CSR of
 * the transpose.  The output is sparse with sorted indices:
 * c_nzind/c_nzval must hold c_len entries; *c_nnz receives the count. */
int gg_spmspv(gg_ctx* ctx, int32_t* c_nzind, void* c_nzval, int64_t* c_nnz, int c_dtype,
              int64_t c_len, const gg_matrix* A, const int32_t* b_nzind, const void* b_nzval,
              int64_t b_nnz, int b_dtype, int64_t b_len, int add, int mul,
              const void* mask, int mask_dtype, int64_t mask_len, int mask_complement);

/* ------------------------------------------------------------- algorithms */
/* Level-synchronous BFS over A_T (row v = in-neighbours of v); dist[v] = level,
 * unreached = typemax(Int32); source is 1-based.  src/algorithms/bfs.jl:1-111. */
int gg_bfs_distances(gg_ctx* ctx, const gg_matrix* A_T, int64_t source, int32_t* dist,
                     int use_mask);
/* gg_bfs_distances without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as gg_bfs_distances; the traversal state (3
 * bitmaps and the level counters, about 24 MiB) lives in a fixed device array
 * of the library, one per GPU, held until the process ends, so the state
 * rows (A_T->m, or the relabelled SELL's internal ids) must not exceed 2^26
 * (ArgumentError otherwise).  Levels are host-stepped (no captured graphs).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (two host threads, or two streams) race and must be serialised by the
 * caller. */
int gg_bfs_distances_noalloc(gg_ctx* ctx, const gg_matrix* A_T, int64_t source, int32_t* dist,
                             int use_mask);
/* BFS parents: parents[source] = 0, unreached = typemax(Int32), otherwise the
 * 1-based frontier in-neighbour chosen by GPUGraphs_any (CSR: first in row
 * order; SELL: largest).  src/algorithms/bfs.jl:113-183. */
int gg_bfs_parents(gg_ctx* ctx, const gg_matrix* A_T, int64_t source, int32_t* parents,
                   int64_t batch_size);
/* Direction-optimising BFS levels: pull (masked SpMV over A_T) while the
 * frontier is large, push (SpMSpV over A's out-edges) while it is small.
 * A_out is CSR of A (row u = out-neighbours of u); for a symmetric graph pass
 * A_T again.  alpha/beta < 0 pick the tuned switch thresholds. */
int gg_bfs_distances_do(gg_ctx* ctx, const gg_matrix* A_T, const gg_matrix* A_out,
                        int64_t source, int32_t* dist, double alpha, double beta);
/* Jacobi Bellman-Ford (min, safe_add) from one source; dist has dist_dtype
 * (Int32 for Bool matrices, else the value type), src/algorithms/shortest_path.jl:12-83. */
int gg_shortest_path(gg_ctx* ctx, const gg_matrix* A_T, int64_t source, void* dist,
                     int dist_dtype, int64_t* iterations);
/* k-source Bellman-Ford on an m x k column-major dist,
 * src/algorithms/shortest_path.jl:88-161.  sources is a host array. */
int gg_shortest_path_multi(gg_ctx* ctx, const gg_matrix* A_T, const int32_t* sources,
                           int64_t k, void* dist, int dist_dtype, int64_t* iterations);
/* The distances of gg_shortest_path by delta-stepping, the bucketed order of
 * the relaxations that src/algorithms/shortest_path.jl:49-53 prepares (its
 * `updated` mask) and leaves unused: vertices below a threshold are expanded
 * push-style over A (CSR of A: row u = out-neighbours of u with the weights
 * of u -> v; for a symmetric matrix, A_T itself) until none is left below it,
 * then the threshold moves by delta to the next bucket that holds a pending
 * vertex.  delta = 0 takes 4 x the mean weight.  The weights reached from
 * the source must be non-negative (a negative or NaN one is GG_ERR_ARGUMENT).
 * phases (optional) receives the expansion phases run. */
int gg_shortest_path_delta(gg_ctx* ctx, const gg_matrix* A, int64_t source, void* dist,
                           int dist_dtype, double delta, int64_t* phases);
/* k delta-stepping runs (1 <= k <= 1024, k * m < 2^32) that share their
 * phases: the expansions of every source run in the same launches, so a
 * phase's fixed cost is paid once for all k.  sources (host, 1-based) and
 * the m x k column-major dist as in gg_shortest_path_multi; column j equals
 * gg_shortest_path_delta from sources[j].  The batched form of the reference's
 * k-source shortest_path (src/algorithms/shortest_path.jl:88-161). */
int gg_shortest_path_delta_multi(gg_ctx* ctx, const gg_matrix* A, const int32_t* sources,
                                 int64_t k, void* dist, int dist_dtype, double delta,
                                 int64_t* phases);

/* ------------------------------------------------------------ batched form */
/* Multi-source BFS as or-and SpMM over 64-bit source bitmasks: k sources
 * (any k >= 1, ceil(k/64) words per vertex).  dist is m x k column-major
 * (dist[:, s] equals gg_bfs_distances from sources[s]). */
int gg_msbfs_distances(gg_ctx* ctx, const gg_matrix* A_T, const int32_t* sources, int64_t k,
                       int32_t* dist, int64_t* levels);
/* Multi-source min-plus SSSP over a row-major m x k working set; dist is
 * m x k column-major like gg_shortest_path_multi. */
int gg_mssp_distances(gg_ctx* ctx, const gg_matrix* A_T, const int32_t* sources, int64_t k,
                      void* dist, int dist_dtype, int64_t* iterations);
/* One BFS per packed graph (gg_pack_graphs layout; A_T the block-diagonal
 * CSR, offsets and sources host arrays, sources 1-based within each graph):
 * the bottom-up levels run in ONE cooperative launch of a persistent kernel
 * fed by an atomic chunk queue; a symmetric unit pattern (GG_FLAG_SYMMETRIC)
 * first runs top-down levels over a queue of the claimed rows while their
 * degree sum * the context's push threshold is below the entries not yet
 * reached (0 = never).  dist[offsets[g] + v] equals gg_bfs_distances(graph g,
 * sources[g])[v]; levels (host, optional) receives each graph's level count.
 * The reference runs bfs_distances (src/algorithms/bfs.jl:1-111) once per
 * graph. */
int gg_bfs_packed(gg_ctx* ctx, const gg_matrix* A_T, const int64_t* offsets, int n_graphs,
                  const int32_t* sources, int32_t* dist, int32_t* levels);
/* The same results as gg_bfs_packed from ONE direction-optimising traversal
 * of the packed graph seeded with every graph's source at level 0 (the blocks
 * share no edge, so each vertex's level is its level within its own graph):
 * push levels over A_out's rows while the frontier is small, pull levels over
 * A_T (any CSR or SELL layout, relabelled included) while it is large, as in
 * gg_bfs_distances_do (alpha, beta <= 0: the defaults).  A_out is the CSR of
 * the packed A (for a symmetric pack, the packed CSR itself); offsets
 * (n_graphs + 1, rising from 0 to m, no padding needed) and sources (1-based
 * within each graph) are host arrays; levels (host, optional) receives the
 * largest level reached in each graph.  No counterpart in the reference,
 * which runs bfs_distances (src/algorithms/bfs.jl:1-111) once per graph. */
int gg_bfs_packed_do(gg_ctx* ctx, const gg_matrix* A_T, const gg_matrix* A_out,
                     const int64_t* offsets, int n_graphs, const int32_t* sources, int32_t* dist,
                     int32_t* levels, double alpha, double beta);

/* ------------------------------------------------------------------ tuning */
/* The autotuner (no counterpart in the reference): candidate layouts of one
 * graph are built on the device from its CSR and timed on a workload; the
 * fastest is returned.  A gg_tune_candidate is a CSR with a lane count, or a SELL of
 * the gg_sell_plan_device parameters; ms receives the measured time (best of
 * reps, -1 when the candidate does not apply, e.g. relabelling a non-square
 * matrix).  Workloads: one plus-times fp32 SpMV (other value types are timed
 * on fp32 ones); BFS levels from every root; one multi-source BFS of up to 64
 * roots; Bellman-Ford from the first root; Bellman-Ford from up to 32 roots
 * at once.  roots: host, 1-based. */
typedef enum gg_tune_workload {
  GG_TUNE_SPMV = 1, GG_TUNE_BFS = 2, GG_TUNE_MSBFS = 3, GG_TUNE_SSSP = 4, GG_TUNE_MSSSP = 5
} gg_tune_workload;
typedef struct gg_tune_candidate {
  int32_t format;      /* GG_CSR or GG_SELL */
  int32_t csr_lanes;   /* CSR: lanes per row (0 = by mean row length) */
  int64_t slice_size, sigma, heavy_threshold, heavy_seg_len;  /* SELL */
  int32_t sell_flags;  /* GG_SELL_RELABEL | GG_SELL_SORT_ENTRIES */
  int32_t reserved;
  double ms;
} gg_tune_candidate;
/* The default candidate set for a CSR A (max_degree: its longest row):
 * CSR with 1, 4, 16 and 64 lanes; relabelled SELL-64 with heavy thresholds
 * 256 / 1024 / 4096 below the longest row and segments of 4096 / 65536, with
 * and without sorted row entries; SELL-32; SELL-64 with sigma 4096 windows.
 * *n_out receives the count (out must hold capacity entries). */
int gg_tune_default_candidates(const gg_matrix* A, int64_t max_degree, gg_tune_candidate* out,
                               int64_t capacity, int64_t* n_out);
/* Times every candidate (their ms is written) and copies the fastest to *best;
 * the context's CSR lane count is left as it was. */
int gg_tune_layout(gg_ctx* ctx, const gg_matrix* A_csr, int workload, const int32_t* roots,
                   int64_t n_roots, gg_tune_candidate* candidates, int64_t n_candidates, int reps,
                   gg_tune_candidate* best);
/* Times gg_msbfs_distances (k <= 64 roots) under each direction switch alpha
 * (gg_ctx_set_push_threshold; 0 = always bottom-up), writes ms[i] when ms is
 * given, leaves the fastest set on the context and in *best_alpha. */
int gg_tune_push_threshold(gg_ctx* ctx, const gg_matrix* A_T, const int32_t* roots, int64_t k,
                           const double* alphas, double* ms, int64_t n_alphas, int reps,
                           double* best_alpha);
/* The same for gg_shortest_path_multi (k <= 32 roots; A_T with its weights)
 * under each sparse-iteration alpha (gg_ctx_set_sssp_threshold). */
int gg_tune_sssp_threshold(gg_ctx* ctx, const gg_matrix* A_T, const int32_t* roots, int64_t k,
                           const double* alphas, double* ms, int64_t n_alphas, int reps,
                           double* best_alpha);

/* ---------------------------------------------------------------- serving */
/* Continuous-batching query server over HBM-resident graphs (no counterpart
 * in the reference, which answers one call at a time).  Each graph gets a BFS
 * engine (64 bitmask slots) and an SSSP engine (64 min-plus columns) on
 * streams of their own; an SSSP query on a Bool graph whose values are all
 * true (GG_FLAG_UNIT_PATTERN: every edge weighs one) is served by the BFS
 * engine.  A query joins a free slot between two steps of its engine, its
 * result is identical to gg_bfs_distances / gg_shortest_path from its source, and it is admitted only when an n-element result block
 * fits in the result arena of budget_bytes (else it waits; a query that can
 * never fit fails with GG_ERR_BUDGET).  Graph arrays stay owned by the
 * caller and must outlive the server. */
typedef struct gg_server gg_server;
typedef enum gg_query_kind { GG_QUERY_BFS = 1, GG_QUERY_SSSP = 2 } gg_query_kind;
typedef struct gg_query_info {
  int64_t request_id;
  int32_t graph, kind;
  int64_t source;
  int32_t status;      /* GG_OK, GG_ERR_BUDGET, GG_ERR_NOT_CONVERGED or the launch error */
  double arrival_s;    /* host steady clock, seconds */
  double admit_s;
  double complete_s;
} gg_query_info;
typedef struct gg_server_counters {
  int64_t completed, failed, waiting, running, steps;
  size_t arena_bytes, arena_used;
} gg_server_counters;
int gg_server_create(int device, size_t budget_bytes, gg_server** out);
int gg_server_add_graph(gg_server* s, const gg_matrix* A_T, int* graph_id);
/* Step pacing (no counterpart in the reference): while other engines have
 * steps in flight, an engine whose slots are not all busy waits
 * factor x (its mean step time on the device, start to end of a step on its
 * stream) x (its free slot fraction) x (its step time / the longest engine
 * step time) after its last step before the next one, so that arriving
 * queries share the steps of expensive engines while cheap ones step at
 * once; 0 steps whenever a slot is busy.  Default 24. */
int gg_server_set_pacing(gg_server* s, double factor);
int gg_server_start(gg_server* s);
int gg_server_submit(gg_server* s, int graph_id, int kind, int64_t source, int64_t* request_id);
/* Finished (or failed) queries not reported before; waits up to wait_s for one. */
int gg_server_poll(gg_server* s, gg_query_info* out, int64_t max, int64_t* n_out, double wait_s);
/* Copy a finished query's n-element result (Int32 levels for BFS; the dist
 * type of gg_shortest_path for SSSP) and free its block.  Waits up to wait_s. */
int gg_server_fetch(gg_server* s, int64_t request_id, void* host_dst, size_t bytes, double wait_s);
int gg_server_fetch_device(gg_server* s, int64_t request_id, void* dev_dst, size_t bytes,
                           double wait_s);
int gg_server_release(gg_server* s, int64_t request_id, double wait_s);
int gg_server_stats(gg_server* s, gg_server_counters* out);
/* Per engine of a graph (kind GG_QUERY_BFS: its BFS engine, which also serves
 * the unit-weight SSSP queries; GG_QUERY_SSSP: its Bellman-Ford engine): the
 * steps run, the busy slots summed over those steps, the queries finished
 * and the device seconds of all steps (each timed from its first to its last
 * kernel on the engine's stream; engines on other streams run meanwhile, so
 * the sum over engines can exceed the wall time).  No counterpart in the
 * reference. */
typedef struct gg_engine_counters {
  int64_t steps, busy_slot_steps, finished;
  int32_t slots, reserved;
  double device_s;
} gg_engine_counters;
int gg_server_engine_stats(gg_server* s, int graph_id, int kind, gg_engine_counters* out);
int gg_server_stop(gg_server* s);   /* finish the steps in flight and join */
int gg_server_destroy(gg_server* s);

#ifdef __cplusplus
}
#endif

#endif /* GPUGRAPHS_H */
