/*
 * tensororder_hip.h -- C ABI of the MI355X (gfx950) contraction executor.
 *
 * The reference executes a contraction plan through a Python "tensor library" plugin
 * (src/tensor_network/tensor_apis/base_api.py:8-28): per slice, a post-order walk that
 * calls tensordot on every join (src/tensor_network/tensor_network.pyx:142-156,
 * tensor_apis/numpy_apis.py:42-43), summing slice scalars in FP64 (base_api.py:17-28).
 * This library replaces that arithmetic.  The host compiles a plan into a list of launch
 * records whose descriptors live in device memory owned by the caller; every entry point
 * takes plain pointers and sizes, never throws, and returns a TO_* status code.
 *
 * Tensors are dense all-dimension-2 arrays in C order: a rank-r tensor is a vector of 2^r
 * entries and axis 0 is the most significant bit of the flat index.  A join contracts X
 * and Y over s shared bits:
 *     C[batch][(f << shf) | (y << shy)] =
 *         sum_k X[batch*xb + pdep(f, xf_pos) + pdep(k, xs_pos)]
 *             * Y[batch*yb + pdep(y, yf_pos) + pdep(k, ys_pos)]
 * where pdep(v, pos) places bit i of v at bit position pos[i].  A batch stride of 0
 * broadcasts an operand across the batch (slice-invariant tensors).  Free-bit positions
 * (xf_pos, yf_pos) must increase with the bit index.
 *
 * Entry types: TO_F64 computes in float64 throughout.  TO_F32S computes in float32 and
 * keeps, per tensor and batch item, a power-of-two exponent (to_scale_rec): true value =
 * stored value * 2^exp; consumers rescale each operand by 2^-e, e = frexp exponent of the
 * operand's max |entry|, so values never under- or overflow.  Slice results are summed in
 * float64 as (mantissa, int64 exponent) pairs in both modes.
 */
#ifndef TENSORORDER_HIP_H
#define TENSORORDER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TO_OK 0
#define TO_ERR_OOM 1     /* caller should slice more: tensor_network.OutOfMemoryError */
#define TO_ERR_INVALID 2 /* malformed launch record or descriptor */
#define TO_ERR_HIP 3     /* HIP runtime error; see to_last_error() */

/* entry types */
#define TO_F64 0
#define TO_F32S 1

/* launch record op codes */
#define TO_OP_LEAF_GATHER 1  /* descs: to_leaf_desc[count] */
#define TO_OP_JOIN_STREAM 2  /* descs: to_join_desc[count]; variant = y values per thread (YC),
                                or TO_STREAM_VEC | YC for the 16-byte form: fx >= 10 (float) /
                                9 (double), the lowest 2 / 1 bits of X free, shy == 0,
                                shf == fy, YC == 2^fy, LDS 8*2^s (16-aligned) + YC*2^s entries;
                                or TO_STREAM_ROWS for the row form: shy == 0, shf == fy,
                                5 <= fy <= 10 (float) / 9 (double), fx >= 6, s <= 4, no LDS */
#define TO_STREAM_VEC 0x100
#define TO_STREAM_ROWS 0x200
#define TO_STREAM_ANY 0x400  /* one launch of streamed joins of mixed variants, each given by
                                its descriptor's kvar; lds = the largest any of them needs */
#define TO_STREAM_WAVE 0x800 /* one wave (64 lanes) per join, for small joins (s <= 6, ranks
                                of X, Y and C below 32): descriptor i is run by wave i % 4
                                of block i / 4 (grid >= count / 4, rounded up; block0 /
                                nblocks / kvar unread), no LDS */
#define TO_OP_JOIN_SPLITK 3  /* descs: to_join_desc[count]; partial sums (double) to d.partial */
#define TO_OP_JOIN_FINISH 4  /* descs: to_join_desc[count]; sums d.partial into C; variant 1:
                                one block per output (for nsplit >= 32) */
#define TO_OP_SCALE_CLEAR 5  /* aux0: to_scale_rec*, iaux0: record count */
#define TO_OP_ROOT_REDUCE 6  /* aux0: root values, aux1: to_scale_rec* (F32S) or 0,
                                aux2: to_accum*, aux3: wave params (int64[2]) */
#define TO_OP_WAVE_ADVANCE 7 /* aux3: wave params; adds iaux0 to params[0] */
#define TO_OP_STEP_BEGIN 8   /* aux2: to_accum* (zeroed), aux3: wave params (params[0] = 0) */
#define TO_OP_PERMUTE 9      /* descs: to_perm_desc[count]; variant 1: 16-byte chunk kernel,
                                which needs nu == 13 (float) / 12 (double) and 16-byte
                                aligned buffers and strides; variant 3: the same with
                                nu == 14 / 13 (64 KB tiles); variant 0: nu <= 12 */
#define TO_OP_GEMM 10        /* descs: to_gemm_desc[count]; variant bit 0: 128-row output
                                tiles (mlog >= 7), 128 columns (nlog >= 7) or, with bit 1,
                                64 columns (nlog >= 6), instead of 64x64 (>= 6); bit 1: float
                                partial sums accumulated in double every 4 K-steps; bit 2
                                (with bit 0): the product is formed transposed, so a float
                                output with ldcn == 1 is written with 16-byte stores; bit 3
                                (with bit 0): 128x128 tiles (nlog >= 7) at one block per CU,
                                LDS double-buffered (variants 9, 11, 13, 15); bit 4 (float,
                                with bit 0, 128x128 tiles): the product on the bf16 matrix
                                cores, each operand split into three bf16 parts and six
                                partial products summed (with bit 3: variants 25, 27, 29,
                                31; with bit 1, klog >= 7) */
#define TO_OP_ROOT_STORE 11  /* descs: to_root_desc[count]: one instance's root -> its accumulator */
#define TO_OP_BATCH_BEGIN 12 /* aux2: to_accum[iaux0] zeroed; aux3: int64 start time (device clock) */
#define TO_OP_LEAF_BUILD 13  /* descs: to_leaf_build_desc[count]: leaf tables written from their
                                compact descriptions (OR / variable leaves) */

#define TO_MAX_BITS 32

/* One join of a launch group (a record in device memory). */
typedef struct to_join_desc {
    uint64_t x, y, c;          /* device pointers to X, Y, C (element type per entry type) */
    int64_t xb, yb, cb;        /* batch strides in elements (0 = broadcast) */
    uint64_t xs, ys, cs;       /* to_scale_rec* per batch item (TO_F32S), else 0 */
    int64_t xsb, ysb;          /* scale-record batch strides (0 or 1) */
    uint64_t partial;          /* split-K partial sums, double, nsplit per output
                                  (TO_OP_JOIN_SPLITK / _FINISH) */
    int32_t fx, fy, s;         /* X free bits, Y free bits, shared bits */
    int32_t shf, shy;          /* output bit offset of f and of y */
    int32_t nbatch;            /* batch items */
    int32_t block0, nblocks;   /* blocks of this join within its launch */
    int32_t nych;              /* y chunks (stream) */
    int32_t nsplit;            /* k splits (split-K) */
    int32_t lnych;             /* log2(nych) */
    int32_t stage;             /* stream, fx >= 8, bit 0: outputs staged through LDS (shy == 0,
                                  shf == fy, one y chunk; 256 * 2^fy entries after the Y
                                  slab); bit 1: X tile staged (all shared bits of X below
                                  bit 8 + s; 256 * 2^s entries after that) */
    int32_t kvar;              /* stream: this join's own variant (YC, TO_STREAM_VEC | YC or
                                  TO_STREAM_ROWS) when the launch variant is TO_STREAM_ANY */
    int8_t xf_pos[TO_MAX_BITS];
    int8_t xs_pos[TO_MAX_BITS];
    int8_t yf_pos[TO_MAX_BITS];
    int8_t ys_pos[TO_MAX_BITS];
} to_join_desc;

/* One leaf copied per batch item from its table of slice rows. */
typedef struct to_leaf_desc {
    uint64_t src;          /* table: rows x 2^rank entries */
    uint64_t dst;          /* nbatch x 2^rank entries */
    uint64_t coef;         /* int64 [ngroups]: row = sum_g coef[g] * bit_g(slice) */
    uint64_t rec;          /* to_scale_rec* per batch item (TO_F32S), else 0 */
    int64_t exp0;          /* exponent of every stored entry (TO_F32S) */
    int32_t rank, ngroups; /* ngroups = number of slice groups of the whole plan */
    int32_t nbatch;
    int32_t block0, nblocks;
    int32_t pad;
} to_leaf_desc;

/* One leaf table written on the device from its description
 * (src/tensor_network/tensor_network_constructions.py:44-159).  The table holds the 2^rank
 * entries of the leaf with its axes reordered (sliced axes first, as tensor.py:95-119
 * reorders them): table bit q (from the least significant) is bit pos[q] of the leaf's own
 * C-order index i.  With the bits of i as the axis values (axis a is bit rank-1-a):
 *   TO_LEAF_VAR: rank 0 -> val0; i == 0 -> val0; i == 2^rank - 1 -> val1; else 0;
 *   TO_LEAF_OR:  literal bits l = i ^ negmask (negmask: bits of negative literals);
 *                out_bit < 0: val1 if l != 0 else val0;
 *                else val1 if bit out_bit of l == (l without that bit != 0) else val0.
 * val0 / val1 are the values as stored (the weights in float64; their float32 encoding in
 * TO_F32S). */
#define TO_LEAF_VAR 0
#define TO_LEAF_OR 1
typedef struct to_leaf_build_desc {
    uint64_t dst;          /* 2^rank entries */
    double val0, val1;
    int64_t negmask;
    int32_t kind, rank, out_bit;
    int32_t block0, nblocks;
    int32_t pad;
    int8_t pos[TO_MAX_BITS];
} to_leaf_build_desc;
/* Per-tensor, per-batch-item scale (TO_F32S). */
typedef struct to_scale_rec {
    int64_t exp;      /* true value = stored value * 2^exp */
    uint32_t maxbits; /* float bits of max |stored entry| (atomicMax on the bits) */
    uint32_t pad;
} to_scale_rec;

/* Running FP64 sum of slice results: value = mant * 2^exp. */
typedef struct to_accum {
    double mant;
    int64_t exp;
    int64_t count; /* slice results accumulated */
    int64_t pad;
} to_accum;

/* A bit permutation of a batch of rank-r tensors: input bit p goes to output bit
 * out_pos(p).  The tile is the union of the lowest input and output bits (u: input
 * order, v: output order); the remaining bits enumerate tiles. */
typedef struct to_perm_desc {
    uint64_t in, out;
    int64_t inb, outb;   /* batch strides in elements */
    int32_t rank, nu, nrest, nbatch;
    int32_t block0, nblocks;
    int8_t u_in[16];     /* input bit of tile bit j (u order) */
    int8_t v_out[16];    /* output bit of tile bit j (v order) */
    int8_t v2u[16];      /* tile bit j of v order is tile bit v2u[j] of u order */
    int8_t rest_in[32];  /* input bit of rest bit j */
    int8_t rest_out[32]; /* output bit of rest bit j */
} to_perm_desc;

/* A batched dense product C[m][n] = sum_k A[m][k] * B[n][k]: A is 2^mlog x 2^klog and
 * B is 2^nlog x 2^klog, both row-major with k contiguous; C entry (m, n) is stored at
 * c + batch*cb + m*ldcm + n*ldcn.  Requires mlog, nlog >= 6 (mlog >= 7 and nlog >= 7, or
 * nlog >= 6 for the wide form, with 128-row tiles) and klog >= 4 (>= 5 for float with
 * 128-row tiles). */
typedef struct to_gemm_desc {
    uint64_t a, b, c;
    int64_t ab, bb, cb;        /* batch strides in elements (0 = broadcast) */
    int64_t ldcm, ldcn;
    uint64_t as, bs, cs;       /* to_scale_rec* of the operands and result (TO_F32S) */
    int64_t asb, bsb;          /* scale-record batch strides */
    int32_t mlog, nlog, klog;
    int32_t nbatch;
    int32_t block0, nblocks;
    int32_t lsplit;            /* 64x64 form (variants 0, 2) only: 0, or split K into 2^lsplit
                                  parts (K >> lsplit >= 16) whose double partial sums go to c
                                  as double[nbatch << (mlog + nlog)][2^lsplit], output o =
                                  (batch << (mlog + nlog)) | (m << nlog) | n, for a
                                  TO_OP_JOIN_FINISH launch (fx = mlog, fy = nlog, nsplit =
                                  2^lsplit) to add up into C */
    int32_t pad1;
} to_gemm_desc;

/* The root scalar of one instance of a batch: stored into its accumulator as
 * (mantissa, exponent), with the device clock at that moment in accum.pad. */
typedef struct to_root_desc {
    uint64_t root; /* the rank-0 result */
    uint64_t rec;  /* its to_scale_rec (TO_F32S) or 0 */
    uint64_t acc;  /* to_accum* */
    int32_t block0, nblocks;
} to_root_desc;

/* One kernel launch. */
typedef struct to_launch {
    int32_t op;
    int32_t variant;
    int32_t count;  /* descriptors */
    int32_t grid;   /* blocks (256 threads each) */
    int32_t lds;    /* dynamic LDS bytes */
    int32_t flags;  /* TO_LAUNCH_* (graph capture only; to_run runs records in order) */
    uint64_t descs; /* device pointer to the descriptor array */
    uint64_t aux0, aux1, aux2, aux3; /* per op; for the descriptor-table ops (LEAF_GATHER,
                                        JOIN_STREAM, JOIN_SPLITK, JOIN_FINISH, PERMUTE, GEMM,
                                        LEAF_BUILD) aux1 is 0 or a device int32[grid] map of
                                        each block to its descriptor (the last one whose
                                        block0 is at or below the block), read instead of a
                                        binary search over block0 (TO_STREAM_WAVE launches
                                        ignore it: block and wave fix the descriptor) */
    int64_t iaux0, iaux1;
} to_launch;

/* to_launch.flags, read by to_graph_create: the record runs on a second captured stream
 * (SIDE), and/or the side stream is joined into the main one and forked from it again
 * before the record (JOIN), so SIDE records between two JOINs run beside the main ones. */
#define TO_LAUNCH_SIDE 0x1
#define TO_LAUNCH_JOIN 0x2

/* Library ABI version (bumped when a struct layout changes). */
int to_version(void);

/* Size in bytes of each record type, for host-side layout checks. */
int to_struct_sizes(int64_t* join_desc, int64_t* leaf_desc, int64_t* scale_rec,
                     int64_t* accum, int64_t* launch, int64_t* perm_desc, int64_t* gemm_desc,
                     int64_t* root_desc, int64_t* leaf_build_desc);

/* Last error message of the calling thread ("" when none). */
const char* to_last_error(void);

/* Number of visible HIP devices. */
int to_device_count(int* count);

/* Select the device for subsequent calls on this thread. */
int to_set_device(int device);

/* Launch the records in order on `stream` (hipStream_t, 0 = default stream).
 * Replaces the per-node tensordot calls of one post-order walk
 * (src/tensor_network/tensor_network.pyx:142-156) for a whole batch of slices. */
int to_run(const to_launch* launches, int count, int entry_type, void* stream);

/* As to_run, recording a HIP event before each record and after the last on `stream`;
 * waits for completion and stores each record's device time in ms[i] (count floats). */
int to_run_timed(const to_launch* launches, int count, int entry_type, void* stream, float* ms);

/* Capture the records into a HIP graph once; replay it with to_graph_launch.  Records
 * flagged TO_LAUNCH_SIDE are captured on a second stream (created and destroyed here,
 * default priority) between TO_LAUNCH_JOIN points; the graph ends joined. */
int to_graph_create(const to_launch* launches, int count, int entry_type, void* stream,
                    void** graph_exec);
int to_graph_launch(void* graph_exec, void* stream);
int to_graph_destroy(void* graph_exec);

/* Reorder the axes of a rank-r all-dimension-2 tensor: out axis j = in axis perm[j],
 * as numpy.transpose (used by tensor_apis/jax_apis.py:83-88, 146-162 and by
 * tensor.py:112-117).  Element size 4 or 8 bytes. */
int to_permute(const void* in, void* out, int rank, const int32_t* perm, int elem_bytes,
               int64_t batch, void* stream);

/* numpy.tensordot(a, b, (axes_a, axes_b)) of two all-dimension-2 tensors on the device:
 * the per-join primitive of the reference's post-order walk
 * (src/tensor_network/tensor_network.pyx:152-154 -> tensor_apis/numpy_apis.py:42-43).
 * c receives 2^(rank_a + rank_b - 2 nshared) entries in numpy's axis order (a's unshared
 * axes in order, then b's); ranks <= 30, output rank <= 30; c must not overlap a or b.
 * TO_F64: float64 tensors, the records are ignored.  TO_F32S: float32 tensors with one
 * to_scale_rec each (a_rec, b_rec read; c_rec written).  One streamed join, or, for many
 * shared axes and few outputs, split-K partial sums in double in `workspace` (8-byte
 * aligned; to_tensordot_workspace gives the bytes, 0 when none are needed) and a
 * fixed-order finishing sum.  Asynchronous on `stream`. */
int to_tensordot_workspace(int rank_a, int rank_b, const int32_t* axes_a, const int32_t* axes_b,
                           int nshared, int entry_type, int64_t* bytes);
int to_tensordot(const void* a, const void* b, void* c, int rank_a, int rank_b,
                 const int32_t* axes_a, const int32_t* axes_b, int nshared, int entry_type,
                 const void* a_rec, const void* b_rec, void* c_rec, void* workspace,
                 int64_t workspace_bytes, void* stream);
/* to_tensordot without the workspace arguments, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it has no counterpart in the reference.  It allocates nothing: the split-K
 * partial sums go to a file-scope device array of 8 MiB (1M doubles) that each GPU holds
 * until the process ends; a call needing more (to_tensordot_workspace > 8 MiB) returns
 * TO_ERR_INVALID.  Every call on one GPU shares that GPU's array: calls that overlap on
 * one GPU (other streams, other threads) race and must be serialised by the caller. */
int to_tensordot_noalloc(const void* a, const void* b, void* c, int rank_a, int rank_b,
                         const int32_t* axes_a, const int32_t* axes_b, int nshared, int entry_type,
                         const void* a_rec, const void* b_rec, void* c_rec, void* stream);
/* Frequency of the device clock stamped by TO_OP_ROOT_STORE / TO_OP_BATCH_BEGIN, in kHz. */
int to_wall_clock_khz(int* khz);

/* Block until all work on `stream` is done. */
int to_synchronize(void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TENSORORDER_HIP_H */
