// Batched engine: many independent grids (2D or 3D, uneven sizes) advanced
// by one set of launches. Every launch covers the tiles of all grids of the
// batch; a workgroup finds its grid in the launch's entry table (sorted
// block ranges), so the per-element arithmetic, the tile decomposition of a
// grid and the order of its reductions are those of the grid alone: a
// grid's result does not depend on what else is in the batch.
//
// Used by the static batch (pano_batch_*, hipGraph-captured steps, tile
// shapes from an autotuner) and by the serving loop (pano_server_*, dynamic
// membership).
#pragma once

#include <string>
#include <vector>

#include "mg.hpp"
#include "mg_device.hpp"
#include "mg_rr.hpp"
#include "mg_tb.hpp"
#include "solver_device.hpp"

namespace pano {

// One grid's entry in a batched launch.
struct BEntry {
    Dims D;          // dims at the launch's level (bc included)
    Dims Dc;         // coarse dims (restriction / prolongation)
    Grid3 G;         // this grid's tile grid for the launch (G.nb blocks)
    int start;       // first global block of this grid in the launch
    int sidx;        // index of the grid's solver state
    int64_t o[4];    // byte offsets in the arena of the fields the launch touches
    double s;        // dt, or the level scale dt/4^l
    float rinv[8];   // smoother reciprocals 1/(s*nc) at this level
    int zc;          // one-pass smoothing tiles: planes per chunk (G: tiles x, y, chunks)
};

// One grid's coarsest-level solve (fast diagonalisation).
struct BFdm {
    int start;       // first global block
    int nblk;        // blocks of this grid (nz planes or ny rows)
    int sidx;
    int is3d;        // 3D grid (its coarsest level may have a single plane)
    int nx, ny, nz;
    const double *Vx, *Vy, *Vz, *lx, *ly, *lz;
    double s;
    int64_t o_r, o_z, o_t0;  // arena byte offsets: coarse rhs (f32), solution (f32), FP64 temp
};

// Device view of a launch table: entries sorted by first block, and their
// first blocks in a compact array for the lookup.
struct BTab {
    const BEntry* e;
    const int* start;
    int n;    // entries (at most kMaxEntries)
    int nbt;  // blocks of the launch
};
struct FTab {
    const BFdm* e;
    const int* start;
    int n;
    int nbt;
};
constexpr int kMaxEntries = 4096;

// Host description of one grid of the batch and where its fields live.
struct HGrid {
    Dims D;
    double dt;
    int levels;
    int sidx;          // state index (slot)
    int64_t cells;     // level-0 cell count
    int64_t faces;
    // arena byte offsets (f32 fields unless noted)
    int64_t o_rho[2], o_u[2];
    int64_t o_x, o_r, o_z, o_p, o_res;
    int64_t o_zl[16], o_rl[16];
    int64_t o_wl[16];  // levels above the coarsest: the down leg's z (pre-smoothing output)
    int64_t o_t0;      // FP64 coarse temp
    int64_t bytes;     // total bytes of this grid
    const double *Vx, *Vy, *Vz, *lx, *ly, *lz;  // coarsest eigenbases
};

// Tile shapes of the three launch families.
struct TileShapes {
    int nat_x = 64, nat_y = 4;  // natural-layout cell and face tiles (threads along x / y)
    int rb_x = 64, rb_y = 4;    // red-black x-pair tiles (threads along x / y)
    int tb2 = 0;                // one-pass smoothing tiles of 2D levels (with_tb2 index)
};

// Byte size of a grid's fields and their offsets from `base`. A projection
// grid (proj) has one velocity buffer and no density.
pano_status hgrid_layout(Ctx* c, HGrid& g, int64_t base, bool proj = false);

// Checks a grid for the batched solver (f32, multigrid with `levels`).
pano_status hgrid_check(const Dims& D, int levels);

class BatchCore {
public:
    Ctx* c = nullptr;
    char* arena = nullptr;     // base of all grid fields
    TileShapes shapes;
    int nu = 2;
    int max_iterations = 500;
    double tol = 1e-6;
    SolverState* states = nullptr;  // device, one per slot
    SolveCfg* cfgs = nullptr;       // device, one per slot
    int nslots = 0;
    bool owns_slots = true;         // false: states/cfgs belong to another core

    BatchCore() = default;
    BatchCore(const BatchCore&) = delete;
    BatchCore& operator=(const BatchCore&) = delete;
    ~BatchCore();

    // Builds the launch tables of the grid subset `ids` (indices into `grids`).
    pano_status build(const std::vector<HGrid>& grids, const std::vector<int>& ids);
    // Enqueue the phases for the grids the tables were built for.
    pano_status enqueue_advect(int ping);         // density and velocity advection (2D/3D, no source)
    pano_status enqueue_divergence(int ping);     // r = -div u (u = velocity buffer `ping`)
    pano_status enqueue_init();                   // CG init: x = 0, stats, threshold, z = M r, p = z, sigma
    pano_status enqueue_iteration();              // one CG iteration
    pano_status enqueue_project(int ping);        // u += dt d0*(x), closed walls zeroed
    pano_status ensure_slots(int n);
    // Level-0 red-black sweeps without solver states (tile-shape timing only).
    pano_status tune_smooth();
    // Level-0 one-pass pre- and post-smoothing without solver states (timing only).
    pano_status tune_smooth_tb();

private:
    // A launch table lives in the core's one device table buffer: its
    // entries, then their first blocks (int).
    struct Table {
        std::vector<BEntry> h;
        const BEntry* d = nullptr;
        int nbt = 0;  // total blocks
        BTab view() const {
            return BTab{d, reinterpret_cast<const int*>(d + h.size()), (int)h.size(), nbt};
        }
    };
    struct FdmTable {
        std::vector<BFdm> h;
        const BFdm* d = nullptr;
        int nbt = 0;
        FTab view() const {
            return FTab{d, reinterpret_cast<const int*>(d + h.size()), (int)h.size(), nbt};
        }
    };
    Table nat_;            // level-0 cells, natural tiles (advection)
    Table fx_, fy_, fz_;   // face tiles per component (advection)
    Table div_[2];         // natural tiles: o[0] = r, o[1] = u[ping]
    Table prc_[2];         // natural tiles of the projection: o[0] = u[ping], o[1] = x
    Table flat_;           // flat vector blocks (2048 cells per block)
    Table rb_[16];         // red-black pair tiles per level
    Table ap_;             // level-0 operator tiles: rb_[0]'s, with runs of kApplyRunY rows on 2D grids
    Table rbs_[16];        // the same, for the grids smoothed by half-sweeps at the level
    Table tb_[16][kTbConfigs];  // one-pass smoothing tiles (3D levels, by configuration)
    Table dot_;            // per grid: where the last post-smoothing left its z.r partials
    Table co_[16];         // coarse-cell tiles of the restriction into level l+1
    Table rsd_[16];        // pair tiles of the residual (the grids that restrict separately)
    Table rr2_[16];        // 2D levels: residual + restriction in one pass, coarse-cell tiles
                           // (o[0] = z_l, o[1] = r_l, o[2] = r_{l+1})
    Table prl_[16];        // pair tiles of level l for the prolongation from l+1
    FdmTable plane_, zcol_;  // 3D coarsest solve
    FdmTable rows2_, cols2_;  // 2D coarsest solve
    int levels_ = 0;
    bool any3d_ = false;
    bool split_first_[16] = {};  // first sweep as red cells + black half-sweep at this level
    bool tb_enabled_ = true;      // one-pass smoothing on 3D levels (PANO_MG_FUSED=0 turns it off)
    // the prolongation inside the one-pass post-smoothing (PANO_MG_PRO=0 turns
    // it off): those grids' pre-smoothing writes o_wl[l], which the residual
    // and the post-smoothing read, and they have no prolongation launch
    bool pro_ = true;
    int tb_base_[kTbConfigs] = {};  // first partial slot of each configuration's level-0 launch
    size_t fdm_lds_ = 0;
    size_t rows2_lds_ = 0, cols2_lds_ = 0;  // dynamic LDS of the 2D coarsest passes
    double* part0_ = nullptr;
    double* part1_ = nullptr;
    double* part2_ = nullptr;
    size_t part_cap_ = 0;
    size_t slot_cap_[2] = {0, 0};
    char* tab_dev_ = nullptr;      // all launch tables (device)
    char* tab_host_ = nullptr;     // their pinned staging copy
    size_t tab_cap_ = 0, tab_host_cap_ = 0, tab_used_ = 0;
    size_t part_bytes_[3] = {0, 0, 0};

    pano_status place(Table& t);
    pano_status place(FdmTable& t);
    pano_status ensure_partials(size_t n);
    void fin(int op, const double* p0, const double* p1, const Table& t, int use_max);
    pano_status vcycle(bool with_dot);
    void smooth_tb_level(int l, bool post, bool dot, const SolverState* st);
};

}  // namespace pano
