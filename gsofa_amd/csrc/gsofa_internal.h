// gsofa_internal.h -- structures shared by the host driver and the CDNA4
// kernels.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "gsofa.h"

struct gsofa_plan;

namespace gsofa {

// Device control block.  Zeroed by the host before every traversal launch.
struct Ctrl {
    // generation barrier
    uint32_t bar_count;
    uint32_t bar_gen;
    uint32_t abort;        // set when a bounded device wait gives up
    uint32_t pad0;
    // rotating per-level counters (level L reads slot L%3)
    unsigned long long qcnt[3];   // queue sizes (pairs appended, may exceed capacity)
    uint32_t cursor[4];           // work-fetch cursors
    unsigned long long orphans;   // frontier pairs parked in the overflow bitmap
    unsigned long long fill_pos;  // fill append position
    // statistics, accumulated once per block at kernel end
    unsigned long long fill_count;
    unsigned long long te;
    unsigned long long reach;
    unsigned long long edge_checks;
    unsigned long long levels;
    unsigned long long rescans;
    uint32_t src_next;            // next share-local source index (block mode)
    uint32_t pad1;
    // decision inputs published by the last block to arrive at a grid barrier
    // (slot gen & 1): queue size of the next phase and outstanding orphans
    unsigned long long snap_q[2];
    unsigned long long snap_orph[2];
    // reach mode: TE and |R| of the A^T pass, and frontier vertices expanded
    // (both passes)
    unsigned long long te_t;
    unsigned long long reach_t;
    unsigned long long expansions;
};

constexpr int32_t kMaxN = (1 << 25) - 1;      // wave-merge totals stay below 2^31
constexpr int64_t kMaxBlockNnz = 1ll << 30;  // block kernel: column byte offsets below 2^32

// GSOFA_OK, or the status gsofa_plan_create would give for A's shape and
// index ranges (n, nnz limits; every column index in [0, n)).
int validate_matrix(const gsofa_matrix* A);

// Per-row upper bounds of |L_i| and |U_i| (strict; A's own entries included)
// from the elimination tree of A + A^T (gsofa_bound.cpp).  rp/ci are packed
// rows, sorted and duplicate-free.
void lu_row_bounds(int32_t n, const int32_t* rp, const int32_t* ci, std::vector<int64_t>& Lb,
                   std::vector<int64_t>& Ub);

constexpr int kBlockSlots = 32;      // largest source-slot count per workgroup in block mode
constexpr int kBatchSlots = 16;      // slots per workgroup for batches of matrices by default

// Block-resident mode: each workgroup owns S source slots (32, 16 or 8).
// Work comes as a table of chunks, processed in table order (decreasing
// sources): chunk c covers sources first, first - stride, ... (count <= S) of
// matrix `mat`, whose vertices are [mat_base[mat], mat_base[mat+1]) in the
// global numbering of a block-diagonal batch (a single matrix is mat 0).
struct BlkChunk {
    uint32_t first;
    uint32_t count;
    uint32_t stride;
    uint32_t mat;
};

struct BlkParams {
    const int32_t* rp;
    const int32_t* ci;
    int32_t n;             // total vertices (all matrices of the batch)
    uint32_t W;            // per-workgroup window: largest matrix size
    const BlkChunk* chunks;
    uint32_t n_chunks;
    const uint32_t* mat_base;  // n_mats + 1 vertex offsets
    uint32_t nblocks;
    uint32_t* lab;         // nblocks * W * S labels, [vertex][slot]
    uint32_t* mask;        // nblocks * 2 * W frontier slot masks
    uint4* qg;             // nblocks * 2 * W frontier entries (v, row begin, row end, -)
    uint64_t* fills;
    uint64_t fill_cap;     // 0: count fills without writing the fill stream
    Ctrl* ctrl;
    uint32_t* bstate;      // per workgroup: max-id offset, groups since re-init
    unsigned long long* row_te;     // n: TE of each processed source row
    unsigned long long* row_reach;  // n: |R_<(s)| of each processed source row
    unsigned long long* row_fills;  // n: fills of each processed source row
    uint32_t loops;        // mini-groups per workgroup between label re-initialisations
};
int blk_block_size();
bool blk_slots_valid(int slots);
int choose_slots(const gsofa_options& o, int32_t n_mats, uint64_t nsrc, int* slots);
// Block-mode sizing shared by plan creation, footprints and the server.
int64_t block_bytes(uint32_t nb, uint32_t W, int slots);
// cols: L + U column entries reserved (sorted + staging, 8 B each); < 0 means
// nnz + fill_cap.
int64_t block_fixed_bytes(int64_t n, int64_t nnz, int32_t n_mats, uint64_t nchunks, uint64_t fill_cap,
                          int64_t cols = -1);
uint64_t default_fill_cap(int64_t nnz);
int block_grid(const gsofa_options& o, int slots, uint64_t nchunks, uint32_t W, int64_t avail, uint32_t* nb_out);
int resident_blocks(int slots, int device, uint32_t* out);
const void* blk_kernel_ptr(int slots);
hipError_t launch_traverse_blk(const BlkParams& p, int slots, size_t lds_bytes, hipStream_t s);

// Reach mode (gsofa_reach.hip): one launch over a work table whose chunks
// run over A (U rows, TE, |R|) or, with rch_pass_bit() set in their stride,
// over A^T (L columns, written as (row, column) fills).
struct RchParams {
    const int32_t* rp;     // packed row pointers and columns of A
    const int32_t* ci;
    const int32_t* rpT;    // of A^T
    const int32_t* ciT;
    uint32_t W;            // per-workgroup window: largest matrix size
    const BlkChunk* chunks;
    uint32_t n_chunks;
    const uint32_t* mat_base;
    uint32_t nblocks;
    void* vis;             // nblocks * W visited words (8 B at 32 slots, 16 B at 64)
    void* nmask;           // nblocks * 2 * W slot masks, this level's and the next's (slots/8 B each)
    uint4* q;              // nblocks * 2 * W frontier entries (v, row begin, row end, -)
    uint64_t* fills;       // entries a workgroup reserved but did not use hold ~0 (set before launch)
    uint64_t fill_cap;     // 0: count fills per row without writing the fill stream
    uint32_t fill_blk;     // fill-stream entries a workgroup reserves at once (rch_fill_block)
    Ctrl* ctrl;
    uint32_t* bstate;      // per workgroup: stamp of its last mini-group
    unsigned long long* row_fills;  // counts only: fills per row, added by both passes (zeroed first)
    unsigned long long* mat_te;     // per matrix TE and |R| of the A pass (zeroed first), or null
    unsigned long long* mat_reach;
    uint32_t loops;        // stamps per workgroup before its visited words are cleared
};
bool rch_threads_valid(int threads);  // threads per reach workgroup: 1024 or 512
bool rch_bits_valid(int bits);
const void* rch_kernel_ptr(int bits, int threads);
int64_t reach_bytes(uint32_t nb, uint32_t W, int bits);
uint32_t rch_pass_bit();
// Fill-stream room of a reach-mode run whose fills are at most `need`: its
// workgroups reserve the stream in blocks and may leave part of each unused
// (at most a sixteenth of a block, and the last block of each workgroup).
// rch_fill_room holds for any grid of up to 1024 workgroups whose block size
// comes from rch_fill_block(need, nblocks).
uint64_t rch_fill_room(uint64_t need);
int64_t rch_visit_bytes(int bits);  // visited word of a reach workgroup: 8, 16 or 32 B
uint32_t rch_fill_block(uint64_t need, uint32_t nblocks);
hipError_t launch_reach(const RchParams& p, int bits, int threads, hipStream_t s);
int reach_resident(int device, int threads, uint32_t* out);
int reach_threads(const gsofa_options& o);  // 0 -> 1024
int resolve_mode(const gsofa_options& o);
// Tuned knobs (gsofa_amd/tune.py's table) for a workload: its signature, and
// the entry applied to options left at their defaults (true if one applied).
std::string tune_signature(int64_t n_max, int64_t nnz, int64_t n_total, int32_t n_mats);
bool apply_tuned(gsofa_options* o, int64_t n_max, int64_t nnz, int64_t n_total, int32_t n_mats);
int choose_bits(const gsofa_options& o, int32_t n_mats, uint64_t nsrc, int* bits);
int reach_grid(const gsofa_options& o, int bits, uint64_t nchunks, uint32_t W, int64_t avail, uint32_t* nb_out);
void transpose_rows(int32_t n, const int32_t* rp, const int32_t* ci, std::vector<int32_t>& rpT,
                    std::vector<int32_t>& ciT);

struct TravParams {
    const int32_t* rp;     // n+1 packed row pointers
    const int32_t* ci;     // nnz column indices, rows contiguous
    int32_t n;
    int32_t n_gpu;         // virtual GPU count (source stride)
    int32_t gpu_id;        // share being processed
    int32_t chunk;         // concurrent sources per group (C)
    int32_t n_groups;
    int32_t first_group;   // groups [first_group, first_group+n_groups) of this launch
    uint32_t* lab;         // packed max-id label arena
    uint64_t lab_words;
    uint32_t* mark;        // C x n fill markers (stamp = source + 1)
    uint64_t* qbuf0;       // frontier pair queues: (slot << 32) | vertex
    uint64_t* qbuf1;
    uint32_t qcap;
    uint32_t orphan_words; // ceil(n/32)
    uint32_t* orphan;      // C x orphan_words overflow bitmap
    uint64_t* fills;       // appended fills: (row << 32) | col
    uint64_t fill_cap;
    Ctrl* ctrl;
    uint32_t nblocks;      // persistent grid size
    int32_t block_threshold; // rows at least this long are expanded by a whole block
    uint32_t off_start;    // max-id offset of the first group of this launch
    uint32_t groups_since_init; // groups already run since the last label re-init
    uint64_t lab_hw;       // label words already initialised before this launch
    uint32_t stamp_base;   // marker stamp of source s is stamp_base + s + 1
    uint32_t loops;        // groups between label re-initialisations
    uint64_t spin_ticks;   // wall-clock ticks a barrier waits before aborting
};

// Launchers (gsofa_kernels.hip)
hipError_t launch_traverse(const TravParams& p, hipStream_t s, int cooperative);
int traverse_block_size();
const void* traverse_kernel_ptr();
size_t traverse_lds_bytes();

hipError_t launch_fill(uint32_t* p, uint64_t count, uint32_t value, hipStream_t s);
hipError_t launch_add(int32_t* p, uint64_t count, int32_t delta, hipStream_t s);  // p[i] += delta

// A matrix prepared on the host for a reach-mode arena (the request server
// packs each request when it is submitted, in the submitting thread): rows
// of A and of A^T packed, sorted and duplicate-free, and the output bounds.
struct PackedMatrix {
    int32_t n = 0;
    int64_t nnz = 0;
    std::vector<int32_t> rp, ci, rpT, ciT;
    uint64_t need_fill = 0, need_L = 0, need_U = 0;
};
int pack_matrix(const gsofa_matrix* A, PackedMatrix* M);
// Arena: a reach-mode plan that takes one batch after another in its own
// buffers (grown to the largest batch seen, never freed between batches).
int arena_create(const gsofa_options* opt, gsofa_plan** out);
// Load a batch (device bytes within `budget`); then gsofa_plan_execute.
int arena_load(gsofa_plan* P, const PackedMatrix* const* mats, int32_t count, int64_t budget);
// After the execute: per-matrix results copied straight from the device.
int arena_fetch(gsofa_plan* P, int32_t index, gsofa_result* out);
int64_t arena_held_bytes(const gsofa_plan* P);
// Device bytes a batch of reach-mode matrices needs apart from its traversal
// workgroups, and per traversal workgroup.
int64_t reach_fixed_bytes(int64_t n, int64_t nnz, int32_t n_mats, uint64_t nchunks, uint64_t fills,
                          uint64_t lu);
struct FinalParams {
    int32_t n;
    const int32_t* rp;
    const int32_t* ci;
    const uint64_t* fills;
    uint64_t n_fills;      // entries of the fill stream, or with n_fills_dev:
    const unsigned long long* n_fills_dev;  // read on the device, at most n_fills
    int32_t n_gpu;
    int32_t gpu_id;        // -1: all rows processed
    int64_t* L_rowptr;     // n+1
    int64_t* U_rowptr;     // n+1
    int32_t* L_col;
    int32_t* U_col;
    int64_t L_cap, U_cap;  // entries allocated in L_col / U_col: no write beyond
    int64_t* row_fill;     // n
    uint32_t* cntL;        // n scratch (counts, then cursors)
    uint32_t* cntU;
    int64_t* scan_tmp;     // scratch for scans
};
hipError_t launch_count(const FinalParams& f, hipStream_t s);
hipError_t launch_scatter(const FinalParams& f, hipStream_t s);
// Sorts each segment [rowptr[r], rowptr[r+1]) of src into dst; segments that
// end beyond cap entries are skipped.
hipError_t launch_sort_segments(const int64_t* rowptr, const int32_t* src, int32_t* dst, int32_t n, int64_t cap,
                                hipStream_t s);

// Inclusive/exclusive scans over int64 (used for row pointers and supernodes).
size_t scan_tmp_words(int64_t n);
hipError_t exclusive_scan_u32_to_i64(const uint32_t* in, int64_t* out, int64_t n,
                                     int64_t* tmp, hipStream_t s);  // out has n+1 entries

struct SuperParams {
    int32_t n;
    const int64_t* L_rowptr;
    const int32_t* L_col;
    const int64_t* U_rowptr;
    int32_t max_size;
    int32_t* supno;        // n
    int32_t* super_start;  // n+1
    int32_t* n_super;      // 1
    int64_t* tmp64;        // scratch: 2n+1 + scan_tmp_words(n)
    uint32_t* flag;        // n scratch
    const uint32_t* pass;  // optional, n: row i-1 is in L_i (then L_col is not read)
};
hipError_t launch_supernodes(const SuperParams& p, hipStream_t s);
hipError_t spill_rows(const int64_t* Lrp, const int32_t* Lcol, const int64_t* Urp, int32_t n, int32_t n_gpu,
                      int32_t gpu_id, uint32_t* lcnt, uint32_t* ucnt, uint32_t* pass, hipStream_t s);
hipError_t transpose_structure(const int64_t* rowptr, const int32_t* col, int64_t nnz, int32_t n,
                               uint32_t* cnt, int64_t* colptr, int64_t* scan_tmp, int32_t* tmp,
                               int32_t* rows, hipStream_t s);
hipError_t supernode_sub_count(const int32_t* sstart, int32_t ns, const int64_t* Urp, uint32_t* cnt,
                               hipStream_t s);
hipError_t supernode_sub_write(const int32_t* sstart, int32_t ns, const int64_t* Urp, const int32_t* Ucol,
                               const int64_t* xsub, int32_t* sub, hipStream_t s);

}  // namespace gsofa
