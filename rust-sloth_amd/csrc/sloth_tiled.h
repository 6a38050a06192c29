// Tiled raster path (internal).
//
// The W x H cell grid of every frame is cut into TW x TH cell tiles (TW even,
// tiles start on even columns).  With W even, z slot ids = y*W + 2x are even,
// so a tile's slots and the two cells each slot writes (id, id+1) never cross
// a tile edge, and a tile can resolve its cells alone.  Frames of odd width
// take the raster's odd instance (described there).
//
//   k_tile_clear   tile cursors <- 0, newline rows <- -1, overflow <- 0,
//                  careful-list length <- 0; after sizing and before
//                  renders of odd-width frames only: k_tile_raster leaves
//                  what it read in this state for the next render
//   k_bin          1, 2 or 4 triangles per thread: setup, classification,
//                  candidate box, newline rows (LDS max per block, one
//                  atomicMax per row), and the triangle index appended to
//                  every tile its box reaches: boxes within 2 x 2 tiles of one
//                  region count into a 16 x 16 LDS window of tiles, with one
//                  global atomicAdd per window tile per block; others (and the
//                  careful list) append directly
//   k_careful      one wavefront per careful triangle: exact box from row
//                  spans, then its tile entries
//   k_tile_raster  one 256-thread workgroup per tile, grid (tile x, tile y,
//                  frame): each list entry's transformed triangle goes to the
//                  LDS row of its lane, its fragments do a 64-bit LDS
//                  atomicMin of packed keys that also name that row, then
//                  every slot re-evaluates its winner from the row its key
//                  names and stores its two cells and depths (coalesced,
//                  streaming stores).  Lists
//                  longer than the 256 LDS rows run in chunks, storing winners
//                  after each chunk.  The image-mode tail cells are written by
//                  the frame's first tile.  When many boxes are big, a variant
//                  has each wavefront walk its lanes' big boxes with all 64
//                  lanes.
//
// Tile lists live in one pool at fixed per-tile offsets, sized from an exact
// counting run of k_bin plus headroom.  If a later run would exceed a tile's
// capacity (the transforms changed), k_bin raises the overflow flag and the
// host re-plans and renders again.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "sloth_device.h"
#include "sloth_host.h"

namespace sloth {

constexpr uint32_t MAX_GROUPS = 8;        // frame groups of a pipelined launch
// Frames on the tiled path hold fewer triangles than this: the raster's
// 64-bit LDS keys carry the triangle index, the pixel's `sub` bit, a stored
// flag and the entry's 8-bit cache row in their low 32 bits.
constexpr uint32_t TILED_MAX_TRIS = (1u << 22) - 1;
constexpr uint32_t PIPE_MIN_FRAMES = 16;  // frames per group at least

struct TiledPlan {
    uint32_t TW = 64, TH = 32;       // tile size in cells (TW even)
    bool user_tile = false;          // set_tile was called (keeps the size on a re-plan)
    // Frames of even / odd width in the job.  Each parity has its raster
    // instance (a launch skips the other parity's frames); odd widths hold a
    // z slot on every cell column, resolve cells from two slots and finish
    // first-column cells in k_tile_edges.
    bool has_even = true, has_odd = false;
    DevBuf d_edges;                  // odd: per tile and row, the first and last column slots' winners
    uint64_t n_tiles = 0;
    uint64_t n_rows = 0;
    bool sized = false;              // pool offsets match the current layout
    uint64_t generation = 0;         // bumped whenever size() runs (device pointers may change)
    std::vector<uint64_t> off;       // host copy of the per-tile pool offsets (n_tiles + 1)
    DevBuf d_cursor, d_off, d_pool, d_nl, d_overflow;
    // triangles whose candidate box k_careful computes: (frame, triangle,
    // screen vertices) in, 32 bytes each, (triangle, box) out; capacity from
    // the counting run plus headroom
    DevBuf d_careful_cnt, d_careful_in, d_careful_box;
    uint32_t careful_cap = 0;
    // Batches whose draws differ in size, or whose frames differ in tile
    // grid, launch flat grids instead of grids padded to the largest draw
    // (k_bin) or tile grid (the raster), whose extra blocks would only exit:
    // k_bin takes one block per (draw, BIN_BLOCK * TPT triangles) from a
    // table per TPT, the raster one workgroup per tile of the range.
    bool flat_bin = false, flat_raster = false;
    DevBuf d_bin_blocks[3];                      // per TPT 1, 2, 4: (draw, first triangle) of each block
    std::vector<uint64_t> bin_block_begin[3];    // per TPT: each draw's first block (draws + 1)
    // flat raster grids: workgroup i of a launch over tiles [t_lo, t_hi) takes
    // entry t_lo + i, a (tile, frame) pair: in tile order, or (launches of the
    // whole batch) with the tiles of more than four chunks of entries first,
    // so that a tile of thousands of entries (a dense mesh in a small frame)
    // does not start last
    DevBuf d_tile_frame, d_tile_heavy;
    // set by size(): the lists hold enough big boxes (more than 128 pixels,
    // one triangle in 512 or more) for the raster variant whose wavefronts
    // walk those boxes together; SLOTH_BIG_WALK=0/1 overrides
    bool big_walk = false;
    // set by size(): k_bin reduces newline rows through a segment tree
    // (one triangle in 256 or more spans more than 16 rows); SLOTH_BIN_TREE
    // =0/1 overrides
    bool bin_tree = false;
    // -1: the raster variant big_walk picks; 0 / 1: without / with the
    // wavefront walk of big boxes (set by the autotuner)
    int walk_force = -1;

    // Sets the tile grid of every frame and allocates per-tile and per-row
    // state.  Pool offsets become invalid until size() runs.
    int plan(std::vector<Frame>& frames, uint64_t rows);

    // Tile size in cells: TW a power of two in 8..512, TH in 4..64.
    int set_tile(uint32_t tw, uint32_t th);
    // The raster variant the next launches use (-1 = the sizing run's choice).
    void set_walk(int mode)
    {
        if (mode != walk_force) ++generation;   // captured graphs hold the old variant
        walk_force = mode;
    }
    // the even-width instance's variant (the odd one has no big-box walk)
    bool walk() const { return has_even && (walk_force >= 0 ? walk_force == 1 : big_walk); }
    // Triangles per thread of k_bin: 1, 2 or 4 (set by the autotuner).
    int bin_tpt = 4;
    void set_bin_tpt(int t)
    {
        if (t != bin_tpt) ++generation;   // captured graphs hold the old launch
        bin_tpt = t;
    }
    // LDS of the even- or odd-width raster instance at the current tile size
    size_t raster_lds(bool odd) const;
    // Both parities' instances the job needs can take the current tile size
    // (LDS, slots per row).
    bool tile_fits() const;

    // Counting run over all frames (synchronous): per-tile list lengths ->
    // pool offsets with headroom.
    int size(const std::vector<Frame>& frames, const std::vector<Draw>& draws, const Frame* d_frames,
             const Draw* d_draws, hipStream_t s);

    int upload(hipStream_t s) { (void)s; return SLOTH_OK; }

    // clear + bin + careful + raster over frames [f_lo, f_hi).  ev (optional)
    // gets 5 events recorded around the 4 stages.  With a second stream s2
    // and join events (and no ev), ranges of at least 2 * PIPE_MIN_FRAMES
    // frames run as a pipeline of frame groups: the bins of one group overlap
    // the raster of the previous one (join needs groups + 1 events).
    int launch(const std::vector<Frame>& frames, const std::vector<Draw>& draws, const Frame* d_frames,
               const Draw* d_draws, const Ramp* d_ramps, uint32_t f_lo, uint32_t f_hi, uint32_t* cells, float* z,
               int32_t* win, bool win_raw, uint32_t n_ramps, hipStream_t s, std::vector<hipEvent_t>* ev,
               hipStream_t s2 = nullptr, const std::vector<hipEvent_t>* join = nullptr);

    // Reads the overflow flags of frames [0, nf) (synchronises the stream).
    int overflowed(uint32_t nf, hipStream_t s, bool& ovf);

    static std::vector<const char*> stage_names() { return {"k_tile_clear", "k_bin", "k_careful", "k_tile_raster"}; }

  private:
    int launch_clear(const std::vector<Frame>& frames, uint32_t f_lo, uint32_t f_hi, hipStream_t s);
    int launch_bin(const std::vector<Frame>& frames, const std::vector<Draw>& draws, const Frame* d_frames,
                   const Draw* d_draws, uint32_t f_lo, uint32_t f_hi, bool count_only, hipStream_t s);
    int launch_careful(const Frame* d_frames, const Draw* d_draws, bool count_only, uint32_t group, hipStream_t s);
    int launch_raster(const std::vector<Frame>& frames, const Frame* d_frames, const Draw* d_draws,
                      const Ramp* d_ramps, uint32_t f_lo, uint32_t f_hi, uint32_t* cells, float* z, int32_t* win,
                      bool win_raw, uint32_t n_ramps, bool reset_careful, hipStream_t s);
};

}  // namespace sloth
