// knc_device.hpp -- data layout in HBM shared by the kernels and the host
// side of the library.
//
// State of one radial block (nr x nq zones, theta fastest) is five FP64
// planes, D | S_r | S_q | tau | M (volume-integrated conserved quantities and
// scalar mass, the SoA form of BlockState, src/state.rs:26-30), each
// nr*nq doubles, contiguous.  Each block owns three such state slots that the
// RK stages rotate through.  Geometry: the face areas of GridGeometry
// (src/mesh.rs:21-29) and the reciprocal cell volumes as per-zone tables,
// evaluated on the host with the reference's own expressions, and 1-D
// factors for the rest -- per-block radial tables and one theta table shared
// by every block (the reference gives all blocks the same polar grid,
// src/mesh.rs:358-379); the cell volume is volA[i] * dcost[j], the
// reference's expression itself.
#pragma once

#include <stdint.h>

namespace knc {

constexpr int NSLOT = 3;
constexpr int MAX_NQ = 1024;

// Radial geometry table of one block (BlockDesc::rgeo), built on the host:
//   rv[nr+1]    vertices
//   volA[nr]    ((2 pi)(r1^3 - r0^3))/3  dV = volA[i] * dcost[j] (mesh.rs:119-123)
//   rinv[nr]    1/sqrt(r0 r1)            1/r at the zone centre (mesh.rs:157-161)
//   dr[nr]      r1 - r0
struct BlockDesc {
    double* slot[NSLOT];
    const double* rgeo;
    const double* zgeo;   // per-zone geometry (zg_* below)
    float* guess;    // c2p guesses log2(p/D), one nr x nq plane per state slot (guess_plane)
    double* prim;    // primitive planes rho | u_r | u_q | p | rho X (5 x nr x nq), written by k_prim
    int32_t nr;
    int32_t nq;
    int32_t left;    // table index of the inner neighbour, -1 at the inner boundary
    int32_t right;   // table index of the outer neighbour, -1 at the outer boundary
    int32_t order;   // radial position in the chain (0 = innermost)
    int32_t index;   // the caller's block index (BlockIndex.0)
};

// Slot pointer by a runtime index without indexing the array (an indexed
// access would force the descriptor into scratch memory).
__host__ __device__ inline double* slot_of(const BlockDesc& b, int s)
{
    return s == 0 ? b.slot[0] : (s == 1 ? b.slot[1] : b.slot[2]);
}

// Guess plane `s` of a block.  A stage reads the plane of its input slot and
// writes the plane of its output slot (KNC_GUESS_PREVIOUS), so no launch reads
// a plane that it also writes.
__host__ __device__ inline float* guess_plane(const BlockDesc& b, int s)
{
    return b.guess + (int64_t)s * b.nr * b.nq;
}

__host__ __device__ inline const double* rg_rv(const BlockDesc& b) { return b.rgeo; }
__host__ __device__ inline const double* rg_volA(const BlockDesc& b) { return b.rgeo + (b.nr + 1); }
__host__ __device__ inline const double* rg_rinv(const BlockDesc& b) { return b.rgeo + (b.nr + 1) + b.nr; }
__host__ __device__ inline const double* rg_dr(const BlockDesc& b) { return b.rgeo + (b.nr + 1) + 2 * b.nr; }
constexpr int64_t rgeo_size(int nr) { return (int64_t)(nr + 1) + 3 * (int64_t)nr; }

// Per-zone geometry of one block (BlockDesc::zgeo), built on the host with the
// expressions of src/mesh.rs:119-134 as they are written there:
//   ar[(nr+1) nq]   radial face areas, face_area of zone (i, j)'s inner face
//   aq[nr (nq+1)]   polar face areas, face_area of zone (i, j)'s lower face
//   rdv[nr nq]      1/dV of cell_volume (dV = volA[i] dcost[j], the same
//                   expression), the y of div_rcp for U/dV
__host__ __device__ inline const double* zg_ar(const BlockDesc& b) { return b.zgeo; }
__host__ __device__ inline const double* zg_aq(const BlockDesc& b) { return b.zgeo + (int64_t)(b.nr + 1) * b.nq; }
__host__ __device__ inline const double* zg_rdv(const BlockDesc& b)
{
    return b.zgeo + (int64_t)(b.nr + 1) * b.nq + (int64_t)b.nr * (b.nq + 1);
}
constexpr int64_t zgeo_size(int nr, int nq)
{
    return (int64_t)(nr + 1) * nq + (int64_t)nr * (nq + 1) + (int64_t)nr * nq;
}

// Polar table shared by every block, nq+1 entries per array (index nq: the
// last vertex alone):
//   dcost = -(cos q1 - cos q0), cot = cos(qc)/sin(qc) at qc = (q0+q1)/2,
//   dq = q1 - q0.
struct ThetaTable {
    const double* dcost;
    const double* cot;
    const double* dq;
    int32_t nq;
};

constexpr int64_t theta_table_size(int nq) { return 3 * ((int64_t)nq + 1); }

enum { BND_FIXED = 0, BND_OUTFLOW = 1 };

// Boundary rows: primitive + scalar, five planes of 2*nq values
// (rho | u_r | u_q | p | s), rows in radial order (inner side: rows -2,-1;
// outer side: rows nr, nr+1).
struct BoundaryDesc {
    int32_t inner_mode;
    int32_t outer_mode;
    const double* inner_rows;
    const double* outer_rows;
};

// One unit of work of a stage launch: rows [i0, i1) x polar columns
// [j0, j1) of one block.
struct WorkItem {
    int32_t block;
    int32_t i0;
    int32_t i1;
    int32_t j0;
    int32_t j1;
    int32_t pad[3];
};

// Error log of a launch sequence; the smallest key is the reported error.
// A key orders failures by (sequence, block order, row, column): sequence 0
// is the time-step pass, 2k - 1 the model ghost rows of the k-th RK stage of
// the run of steps and 2k that stage (saturating at ERR_SEQ_MAX), so the
// first stage that fails is the one reported, as the reference returns from
// the first failing stage.
// The tag is (block order << 16) | row, block order < 2^15, column < 2^11.
constexpr int ERR_LOG_CAP = 256;
constexpr int ERR_COL_BITS = 11;
constexpr int ERR_TAG_BITS = 31;
constexpr unsigned ERR_SEQ_MAX = (1u << (64 - ERR_COL_BITS - ERR_TAG_BITS)) - 1u;
static_assert(MAX_NQ <= (1 << ERR_COL_BITS), "error keys hold the column in ERR_COL_BITS bits");
__host__ __device__ inline unsigned long long err_key(unsigned seq, int tag, int col)
{
    const unsigned long long s = seq < ERR_SEQ_MAX ? seq : ERR_SEQ_MAX;
    return (s << (ERR_TAG_BITS + ERR_COL_BITS)) | ((unsigned long long)(unsigned)tag << ERR_COL_BITS) |
           (unsigned long long)(unsigned)col;
}
__host__ __device__ inline int err_key_order(unsigned long long key) { return (int)((key >> (ERR_COL_BITS + 16)) & 0x7fff); }
__host__ __device__ inline int err_key_row(unsigned long long key) { return (int)((key >> ERR_COL_BITS) & 0xffff); }
__host__ __device__ inline int err_key_col(unsigned long long key) { return (int)(key & ((1u << ERR_COL_BITS) - 1u)); }
struct ErrEntry {
    unsigned long long key;
    int32_t code;
    int32_t pad;
    double value;
    double cons[4];
};
struct ErrLog {
    unsigned long long min_key;
    unsigned int count;
    unsigned int pad;
    ErrEntry entry[ERR_LOG_CAP];
};

// Newton statistics: iteration histogram, fallbacks, floors.
constexpr int HIST_BINS = 64;
struct C2PStats {
    unsigned long long hist[HIST_BINS];
    unsigned long long zones;
    unsigned long long fallbacks;
    unsigned long long floors;
    unsigned long long learned_iters;
};

}  // namespace knc
