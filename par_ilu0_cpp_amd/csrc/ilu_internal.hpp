// Device state behind lljbash::IluSolver (private to the library).
#pragma once

#include "common.hpp"
#include "ilu_solver.hpp"

#include <vector>

namespace pilu {

// Launch-shape choice for one sweep direction.
struct SweepShape {
    int lanes = 8;  // lanes per row (power of two, 1..64)
};

// Apply M^-1 to device vectors on `stream`, optionally gated by a device
// flag (kernels skip their work while *gate == 0).  Exposed for the GMRES
// solver, which runs on the same stream.
void ilu_apply(lljbash::IluSolver::Ext* ext, const double* d_b, double* d_x, const int* gate);

hipStream_t ilu_stream(const lljbash::IluSolver::Ext* ext);

// Throws if a sweep since the last check hit its watchdog.
void check_sweep_errors(lljbash::IluSolver::Ext* ext);

// Lanes-per-row for a given average row length (the sweep and SpMV
// heuristic; the autotuner may override it).
int lanes_for_avg(double avg_len);

// Level schedule of independent diagonal blocks (block_of[row] in
// [0, nblk)): perm = rows sorted by (block, level inside the block, row),
// blev[b] = first (block, level) group of block b (nblk + 1 entries),
// loff[g] = first position of group g.  Levels of the L (upper = false) or
// U dependency graph.
void block_level_schedule(lljbash::IluSolver::Ext* ext, bool upper, const int* block_of, int nblk,
                          DevBuf<int>& perm, DevBuf<int>& blev, DevBuf<int>& loff);

}  // namespace pilu

namespace pilu {
// Largest block of the block-local sweeps (its x lives in LDS).
constexpr int kMaxBlockRows = 16384;

// Schedule-ordered copy of one direction's factors for the packed sweep.
struct SweepPack {
    DevBuf<int> rowinfo;            // (row, entry count) per schedule position
    DevBuf<int> pcol, psrc;         // slab columns (in-block: -(1 + block slot)) and source CSR positions
    DevBuf<int> kptr;               // first schedule position of every processed block
    DevBuf<double> pval64, pdiag64;
    DevBuf<float> pval32, pdiag32;
    std::vector<int> h_blev, h_loff, h_base, h_width;  // host copies for pass lists
    long ne = 0;
    int max_w = 0;
    int version = -1;               // partition the pack was built for
    long factors = -1;              // factorization its values were gathered from
    DevBuf<char> pdesc;             // PassDesc records, each block's list padded
    DevBuf<int> desc_ptr;           // first PassDesc of each block
    int max_np = 0;                 // most PassDesc records of one block
    int pass_g = 0;                 // lanes per row the pass list was built for
};
// Record streams of the wave-specialised sweep (mode 5) for one direction,
// built for a partition and a lanes-per-row count.
struct WsPack {
    DevBuf<uint4> stream;           // records of every block, blocks in processing order
    DevBuf<int> vdst, vsrc;         // value slot (in T units of the stream) <- factor entry (-1: 0)
    DevBuf<int2> imports;           // per import: face slot of the imported row, first record using it
    DevBuf<char> blocks;            // WsBlock per processed block
    DevBuf<double> face;            // face buffer (value-flags of the rows other blocks read)
    long nvals = 0;
    long nstream = 0;               // 16-byte units
    int nface = 0;
    int max_imp = 0;
    int value_bytes = 0;            // 8 or 4: precision the stream was laid out for
    int version = -1;               // partition the stream was built for
    int g = 0;                      // lanes per row it was built for
    long factors = -1;              // factorization its values were gathered from
    bool face_armed = false;        // every face slot holds kPending
    std::vector<int> face_rng;      // per block id: first and end face slot (independent of g)
    int rng_version = -1;           // partition face_rng belongs to
};
}  // namespace pilu

struct lljbash::IluSolver::Ext {
    int n = 0;
    long nnz = 0;
    int max_row = 0;
    int max_lo = 0, max_up = 0;  // most strictly lower / upper entries in a row
    hipStream_t stream = nullptr;

    pilu::DevBuf<int> rp, ci, dg;
    pilu::DevBuf<int> perm_l, perm_u, level;
    int nlev_l = 0, nlev_u = 0;
    pilu::DevBuf<int> lvoff_l, lvoff_u;  // first position of every level in perm_l / perm_u (+ n)
    // per-slot row descriptors of the single-workgroup sweep (mode 4), built
    // on first use for (pattern_version, slots per level)
    pilu::DevBuf<int4> lvdesc_l, lvdesc_u;
    int lvkey_l[2] = {-1, 0}, lvkey_u[2] = {-1, 0};
    pilu::DevBuf<unsigned long long> lvprog;  // [0]: progress word of the mode-4 sweep; [1]: sink; [2]: slot check
    unsigned lv_epoch = 0;
    pilu::DevBuf<int> lvbad;  // per level: 1 = may not use mode 4's LDS cache (scratch of the descriptor build)
    int lvcached_l = 0, lvcached_u = 0;  // levels that use mode 4's LDS cache
    int pattern_version = 0;

    pilu::DevBuf<double> vin;    // values to factorize (scratch)
    pilu::DevBuf<double> fac;    // factors, FP64
    pilu::DevBuf<float> fac32;   // factors, FP32 copy (precision 32)
    pilu::DevBuf<double> y;      // forward-sweep output, kept all-kPending between applies
    pilu::DevBuf<double> hb;     // staging for the host Substitute path
    pilu::DevBuf<int> ctr;       // ticket counters (pairs)
    pilu::DevBuf<int> err;       // error words
    pilu::DevBuf<char> sort_tmp; // radix-sort scratch
    pilu::DevBuf<int> nzmap;     // A-entry -> F-entry map (SymbolicDevice)
    pilu::DevBuf<signed char> fill_level;
    long nnz_a = 0;
    const float* omega_ptr = nullptr;  // per-row compensation (device, caller-owned)
    // Pivot check of the factorization (PivotGuard): 0 none, 1 stop, 2
    // replace; the dcsrilu0 defaults dpar[30] = 1e-16, dpar[31] = 1e-10.
    int pivot_mode = 0;
    double pivot_tiny = 1e-16, pivot_repl = 1e-10;
    int factor_status = 0;  // dcsrilu0-style code of the last factorization

    int precision = 64;
    // Wall-clock limit of every bounded dependency wait of the analysis,
    // factorization and sweep kernels (the symbolic pass gets twice this).
    // A wait past it gives up: the kernel drains and the call throws.
    double watchdog_s = 10.0;
    int lanes_l = 0, lanes_u = 0;  // 0 = heuristic
    // Sweep algorithm per direction: 0 = level-scheduled rows (k_sweep),
    // 1 = block-local over contiguous blocks of blk_rows rows (k_sweep_blk),
    // 2 = packed block-local over the current partition (k_sweep_pp, level
    // barriers), 3 = the same with LDS value-flags, 4 = one workgroup walking
    // the level schedule (k_sweep_lv; no partition), 5 = wave-specialised
    // block sweep over record streams (k_sweep_ws).
    int mode_l = 0, mode_u = 0;
    pilu::SweepPack pack_l, pack_u;
    pilu::WsPack ws_l, ws_u;
    // Block partition of the rows: contiguous blocks of blk_rows rows, or
    // tiles of the grid (tile[] != 0, grid[] declared by SetGridShape).
    int blk_rows = 0, nblk = 0, max_block_rows = 0;
    int tile[3] = {0, 0, 0}, grid[3] = {0, 0, 0};
    int skew[3] = {0, 0, 0};  // tile coordinates u = x + skew0 y + skew1 z, v = y + skew2 z
    std::vector<int> h_block_of;
    // host copies of the pattern (row_ptr, diagonal positions, columns) and
    // of block_of for the wave-specialised stream builder, copied on first
    // use (host_pattern)
    std::vector<int> h_rp, h_dg, h_ci;
    int h_pattern_version = -1, h_part_version = -1;  // what the host copies were taken from
    pilu::DevBuf<int> block_of;
    int part_version = 0;
    long factor_version = 0;
    pilu::DevBuf<int> bperm_l, bperm_u, blev_l, blev_u, loff_l, loff_u;
    // Set while a no-allocation apply runs (SubstituteDeviceNoAlloc): a
    // sweep structure that would have to be built then raises instead, and
    // the level walk keeps its progress words in lvprog_scratch.
    bool no_alloc = false;
    unsigned long long* lvprog_scratch = nullptr;
    bool analysed = false;  // pattern checked and schedules built (SetupMatrix / SymbolicDevice completed)
    bool factored = false;
    bool y_armed = false;
};
