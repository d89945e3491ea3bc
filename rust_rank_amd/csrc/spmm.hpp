// spmm.hpp -- the sliced sparse pass (k_spmm) shared by the single-matrix
// Krylov context and the batched / served tick runner.
//
// One launch runs any number of jobs of the same lane width, value kind and
// column block size.  Each job describes one operand (A or A^T of one matrix)
// in sliced form, the vector it gathers, the vector it produces and where its
// projection partials go; a workgroup map (WgMap) says which workgroup of
// which job each hardware workgroup runs, and on which XCD.
#pragma once
#include "common.hpp"
#include <functional>
#include <vector>

namespace rr {

constexpr uint32_t SEG_FLAG = 0x80000000u;
constexpr uint32_t NO_ROW = 0xFFFFFFFFu;
constexpr uint32_t MAX_COPIES = 32;  // accumulator rows per split row
constexpr uint32_t MAX_PEND = 16;    // segment slices per wavefront stream
constexpr uint32_t NXCD = 8;         // XCDs of an MI355X, each with its own L2

// One operand in sliced form.  A slice is RPW units processed side by side by
// the RPW row groups of one wavefront, SU nonzeros of each unit per column
// block.  Row slices hold RPW rows (length-sorted, so padding is small);
// segment slices hold RPW consecutive chunks of one long row, whose sum is
// one partial of that row.  Each wavefront owns a contiguous stream of slices
// (balanced by longest-processing-time assignment) and of their column
// blocks: block k of a slice is RPW x SU words, group g's SU words contiguous.
struct SpmmArgs {
    const uint32_t *meta;      // per slice: end block in its wavefront's stream | SEG_FLAG
    const uint2 *unit;         // per slice x RPW: {row, scale[row]} (NO_ROW: none); segment slices: {accumulator row, split row}
    const uint32_t *wave_sl;   // per wavefront: first slice (nw + 1)
    const uint32_t *wave_blk;  // per wavefront: first column block (nw + 1); streams end with 2 zero blocks
    const uint32_t *col;       // packed (col | value << 28) or plain column, 0 value = padding
    const uint32_t *val;       // plain values (general matrices), 0 = padding
    const uint32_t *in;        // gathered block vector, stride B
    uint32_t *out;             // produced block vector, stride B
    const uint32_t *scale;     // per-row factor (split rows)
    unsigned long long *lacc;  // [accumulator rows][B] sums of segment partials (zero between passes)
    const uint4 *linfo;        // per split row: {row, first accumulator row, accumulator rows, segments}
    uint32_t *lcnt;            // per split row: segments added in this pass (zero between passes)
    uint64_t gram_off;         // words from the launch's partial base to this job's first partial
    uint32_t in_bytes;         // size of in (below the gathers' out-of-range offset, just under 4 GiB)
    uint32_t p;
    double pd, inv;
    uint32_t b;
};

// Workgroup map of one k_spmm launch.  The hardware deals workgroup h to XCD
// h mod NXCD, so XCD x runs the workgroups k = h / NXCD < xcnt[x] of its own
// list: entries ent[xoff[x] .. xoff[x + 1]), each {k_begin, job, lwg0,
// stride}; workgroup k belongs to the last entry with k_begin <= k and is
// workgroup lwg0 + (k - k_begin) * stride of that job.  Hardware workgroups
// past xcnt[x] exit at once.  Stored as words: the head, then the entries.
struct WgMapHead {
    uint32_t xoff[NXCD + 1];
    uint32_t xcnt[NXCD];
    uint32_t pad[3];
};
constexpr uint32_t WGMAP_HEAD_WORDS = sizeof(WgMapHead) / 4;
static_assert(sizeof(WgMapHead) % 16 == 0, "entries must stay 16-byte aligned");

// Places jobs of nwg[j] workgroups on the XCDs.  A job larger than half of
// an XCD's share of the launch is spread over every XCD (its workgroup w on
// XCD w mod NXCD, as in a launch of that job alone); the others go whole to
// one XCD, largest first onto the least loaded, so the vectors a job gathers
// stay in one XCD's L2.  Returns the map's words; *grid: the hardware
// workgroups to launch.
std::vector<uint32_t> build_wg_map(const std::vector<uint32_t> &nwg, uint32_t *grid);

struct SliceMat {  // one operand (A or A^T) on the device, in sliced form
    uint32_t nrows = 0, nslices = 0, nseg = 0, nlong = 0, nw = 0;
    uint64_t blocks = 0;
    DevBuf meta, unit, wave_sl, wave_blk, col, val, lacc, linfo, lcnt;
};

// A and A^T of a validated CSR matrix with values < p: row pointers and
// column words (packed: column | 4-bit signed value << 28; plain: column, with
// values separately).  The transpose keeps row order within each column
// (src/matrices.rs:734-778).
void build_operands(const rr_csr *A, uint32_t p, bool packed, std::vector<uint32_t> &rp, std::vector<uint32_t> &colw,
                    std::vector<uint32_t> &vals, std::vector<uint32_t> &rpt, std::vector<uint32_t> &colt,
                    std::vector<uint32_t> &valt);

// True when every value is a signed residue in [-8, 7] and m, n < 2^28.
bool csr_packs(const rr_csr *A, uint32_t p);

// Slots (nonzeros plus padding) of both operands' rows in column blocks of su.
uint64_t padded_slots(const std::vector<uint32_t> &rp, const std::vector<uint32_t> &rpt, uint32_t su);
// Column block size (3 or 4) that pads the rows of both operands least.
uint32_t choose_su(const std::vector<uint32_t> &rp, const std::vector<uint32_t> &rpt);

// One operand's CSR arrays on the device: row pointers (nrows + 1), column
// words (packed or plain) and plain values (null when packed).
struct DevOperand {
    const uint32_t *rp, *colw, *vals;
};

// Rows of an operand longer than this are cut into segments (split rows).
uint32_t long_row_threshold(uint64_t nnz, uint64_t nrows, uint32_t B, uint32_t SU, uint32_t nw);

// Lays out one operand for nw wavefronts (nw a multiple of 4); in_rows: rows
// of the vector it gathers.  With dev (an operand with no row above the
// threshold), the host plans the slices from the row lengths in rp alone
// and the device copies the nonzeros from dev's arrays on stream st (colw
// and vals are then not read).
rr_status build_slices(SliceMat &M, uint64_t nrows, const std::vector<uint32_t> &rp,
                       const std::vector<uint32_t> &colw, const std::vector<uint32_t> &vals, bool packed, uint32_t B,
                       uint32_t SU, uint32_t nw, const uint32_t *scale, uint64_t in_rows,
                       const DevOperand *dev = nullptr, hipStream_t st = nullptr);

// Device part of build_slices: the column blocks of row slices sl[s] =
// {first row in the length order, first block}.
void launch_fill_slices(const DevOperand &op, const uint32_t *shortr, uint32_t nshort, const uint2 *sl, uint32_t nsl,
                        uint32_t RPW, uint32_t SU, uint32_t *cw, uint32_t *vw, hipStream_t st);

// Wavefronts to lay the operands out for, given the value kind and column
// block size the matrix gets.
typedef std::function<uint32_t(bool packed, uint32_t su)> NwOf;
struct OperandInfo {
    bool packed = false;     // every value a signed residue in [-8, 7]
    uint32_t su = 4, nw = 0;
    bool on_device = false;  // laid out by the device path
};

// Validates A (row_ptr monotone, columns < n, values < p) and lays out A
// (scaled by row_scale) and A^T (scaled by col_scale2) in sliced form.  The
// entries are uploaded once; the checks, the transpose and the column-block
// streams run on the device on stream st, and the host plans the slices from
// row lengths; a matrix with rows to cut into segments is laid out on the
// host.  su_req: 3 or 4, or 0 for the size that pads least.
rr_status setup_operands(const rr_csr *A, uint32_t p, uint32_t B, uint32_t su_req, const NwOf &nw_of,
                         const uint32_t *row_scale, const uint32_t *col_scale2, SliceMat &sa, SliceMat &sat,
                         OperandInfo *info, hipStream_t st);

// The n x b block vector V (row-major) to the device as n rows of B >= b
// words, zero past column b, after a check on the device that every word is
// below p: the first one that is not is reported as name[index], as a host
// loop over the words would report it.  Synchronises st.
rr_status upload_block_vector(DevBuf &out, const uint32_t *V, uint64_t n, uint32_t b, uint32_t B, uint32_t p,
                              const char *name, hipStream_t st);

SpmmArgs spmm_args(const SliceMat &M, const uint32_t *in, uint64_t in_rows, uint32_t B, uint32_t *out,
                   const uint32_t *scale, uint32_t p, uint32_t b);

// Resident workgroups per CU of the pass kernel for (B, packed, SU, proj).
int spmm_blocks_per_cu(uint32_t B, bool packed, uint32_t su, bool proj = true);

// One launch over the jobs of a workgroup map (device arrays); grid: the
// hardware workgroups build_wg_map gave.  ev0 / ev1 (may be null): stamped
// with the kernel's own start and end.
// proj = false (lane width 64 only): the pass produces its rows and writes no
// projection partials; k_proj64 then takes the projection.
void launch_spmm(uint32_t B, bool packed, uint32_t su, const SpmmArgs *jobs, const uint32_t *wgmap, uint32_t grid,
                 uint32_t *gram_base, hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                 bool proj = true);

}  // namespace rr
