// Internal declarations of the MI355X linear-solve library: device buffers, the row-sliced block layouts the
// kernels stream, level schedules, and the kernel launchers (rsh_kernels.hip) used by the solver drivers
// (rsh_api.cpp) and the host set-up (rsh_setup.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace rsh {

constexpr int kWave = 64;             // CDNA wavefront
constexpr int kSliceRows = 64;        // rows per slice: one row per lane of a wavefront
constexpr int kRedBlocks = 1024;      // partial sums of the deterministic two-pass reductions
constexpr int kMaxNb = 4;             // block sizes built: 1..4
constexpr int kMaxRestart = 256;      // FGMRES restart lengths accepted (2m+1 vectors of n*nb doubles)

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
#define RSH_HIP(x)                                                                                       \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) throw ::rsh::Error(-2, std::string(#x) + ": " + hipGetErrorString(e_));  \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        RSH_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
    }
    void upload(const std::vector<T>& h, hipStream_t s) {
        alloc(h.size());
        if (!h.empty()) RSH_HIP(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void zero(hipStream_t s) {
        if (n) RSH_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Krylov scalars kept on the device; reductions finish with an epilogue that updates them, so one host read
// per iteration suffices.
struct KScalars {
    double rho, alpha, omega, rhoN, rv, sn, tt, ts, rn, bn, beta, tol;
    int state, it, pad0, pad1;
};
// kDoneS: converged on ||s|| with the final update applied (batched solves, where other systems keep iterating)
enum KState : int { kRun = 0, kConvS = 1, kConv = 2, kBreak = 3, kOmegaZero = 4, kDoneS = 5 };
enum Epilogue : int { kEpNone = 0, kEpRho, kEpAlpha, kEpSn, kEpOmega, kEpRn, kEpBn, kEpNorm, kEpRnRho };

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DBuf<double> red;          // kRedBlocks * 4 partial sums
    DBuf<unsigned> ticket;     // last-block counter of the reductions
    DBuf<KScalars> ks;
    DBuf<double> scal;         // general device scalars (dot results, 64 slots)
    KScalars* hks = nullptr;   // pinned host mirror
    double* hscal = nullptr;
    // device time of set-up + solve spans on this stream (HIP events): a span opens at the first set-up or solve call
    // and closes at the end of the solve
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool spanOpen = false;
    double deviceMs = 0;
};

// ------------------------------------------------------------------ row-sliced block layout
// Rows are taken 64 at a time (one per lane of a wavefront).  A slice stores, for each of its `width` slots,
// the column of every lane and the nb*nb block entries element-major: entry e of slot k for lane l lives at
// val[((off + k) * nb2 + e) * 64 + l], so each load instruction of a wavefront reads 64 consecutive values.
// Rows shorter than the slice width are padded with zero blocks pointing at their own row; rows longer than
// the cap are taken out of the slices and handled by one wavefront each (the "long" rows: perforated cells of
// wells after the Schur fill).
struct SliceHost {
    int nslices = 0;
    long long nslots = 0;
    std::vector<int> width, off;   // per slice
    std::vector<int> row, orig;    // per slice*64+lane: working row, original row (-1 = idle lane)
    std::vector<int> col, src;     // per slot*64+lane: working column, source block index (-1 = padding)
    std::vector<short> dcol;       // optional: col - row as 16 bits (idle lanes: 1, i.e. column 0), when all fit
    // long rows
    std::vector<int> lrow, lorig, lptr, lcol, lsrc;
    std::vector<int> dsrc;         // optional per slice*64+lane: diagonal block (ILU backward sweep)
    std::vector<int> ldsrc;        // per long row
    // optional (batched systems): the system owning each slice / long row, -1 when a slice mixes systems
    std::vector<int> owner, lowner;
};

template <typename TA>
struct SliceDev {
    int nslices = 0, nlong = 0, nb = 1;
    DBuf<int2> info;               // (width, slot offset)
    DBuf<int> row, orig, col, src, dsrc;
    DBuf<short> dcol;              // the columns relative to the lane's row (16-bit layout), or empty
    DBuf<TA> val, dval;
    DBuf<int> lrow, lorig, lptr, lcol, lsrc, ldsrc;
    DBuf<TA> lval, ldval;
    DBuf<int> owner, lowner;
    // coarse AMG levels: the rows' sums are formed in G chains folded like the G-lane CSR step (0: one sum per row)
    int chains = 0;
    int lanes = 1;                 // lanes per row of the SpMV slices (laneSliceLayout)
    void release() {
        for (DBuf<int>* b : {&row, &orig, &col, &src, &dsrc, &lrow, &lorig, &lptr, &lcol, &lsrc, &ldsrc, &owner, &lowner}) b->release();
        info.release();
        dcol.release();
        for (DBuf<TA>* b : {&val, &dval, &lval, &ldval}) b->release();
        nslices = nlong = chains = 0;
        lanes = 1;
    }
    size_t bytes() const {
        return info.bytes() + row.bytes() + orig.bytes() + col.bytes() + dcol.bytes() + src.bytes() + dsrc.bytes() + val.bytes() + dval.bytes() +
               lrow.bytes() + lorig.bytes() + lptr.bytes() + lcol.bytes() + lsrc.bytes() + ldsrc.bytes() + lval.bytes() + ldval.bytes();
    }
};

// A level schedule over a slice layout: level l owns slices [sStart[l], sStart[l+1]) and long rows
// [qStart[l], qStart[l+1]).
struct LevelSched {
    std::vector<int> sStart, qStart;
    int levels() const { return int(sStart.size()) - 1; }
};
// consecutive levels grouped for launching: a run of small levels is swept by one workgroup (fused), a large
// level gets its own launch
struct LevelRun { int l0, l1; bool fused; };
// rowsPerLevel (optional) caps the rows of a fused run at maxRunRows
std::vector<LevelRun> levelRuns(const std::vector<int>& workPerLevel, int smallMax, const std::vector<int>* rowsPerLevel = nullptr,
                                int maxRunRows = 0);

// ------------------------------------------------------------------ host set-up (rsh_setup.cpp)
struct HostBsr {
    int n = 0, nb = 1;
    std::vector<int> rp, ci, diag;
};
struct HostCsr {
    int n = 0;
    std::vector<int> rp, ci, diag;
    std::vector<double> v;
};

// levels of the symmetrised lower dependency graph; ord = rows in level order (ascending within a level)
void levelOrder(const HostBsr& A, std::vector<int>& ord, std::vector<int>& start);
// levels of the plain lower (forward) or upper (backward) graph of a scalar CSR (Gauss-Seidel sweeps)
void csrLevels(const HostCsr& A, bool forward, std::vector<int>& ord, std::vector<int>& start);
// greedy distance-1 colouring, largest degree first (colourBlocks, bsr.cpp:599-630)
int colourGraph(const HostBsr& A, std::vector<int>& colour, std::vector<int>& perm, std::vector<int>& start);
// P A P^T pattern with sorted rows and the block map new -> old (permuteBsr, bsr.cpp:632-653)
void permutePattern(const HostBsr& A, const std::vector<int>& perm, HostBsr& P, std::vector<int>& blockMap);
int rowCap(const std::vector<int>& lengths);
// slices over the rows [r0, r1) of working numbering; entries(i) lists (working col, src block) for working row i
struct RowEntries {
    std::vector<int> ptr, col, src;
    std::vector<int> orig;   // working row -> original row
    std::vector<int> dsrc;   // working row -> diagonal block (or empty)
};
// longRow (optional, per working row): rows taken as long rows whatever their length; skipRow (optional): rows left out
void buildSlices(const RowEntries& E, const std::vector<int>& levelStart, int cap, SliceHost& S, LevelSched& L,
                 const std::vector<char>* longRow = nullptr, const std::vector<char>* skipRow = nullptr);
// one-level row slices of a CSR / BSR pattern in natural row order: rows longer than the bulk are long rows, padding
// gathers the lane's own row, and the 16-bit relative columns are kept when every slice row fits them.  coarse (an
// AMG level): the bulk ends a quarter above the 99th percentile of the row lengths, and rows with a column beyond
// 16-bit reach are long rows too
void rowSliceLayout(int n, const std::vector<int>& rp, const std::vector<int>& ci, SliceHost& h, bool coarse = false);
// the same rows in slices of 64 / lanes rows, `lanes` lanes per row (1, 2 or 4) splitting each row's entries
void laneSliceLayout(int n, const std::vector<int>& rp, const std::vector<int>& ci, int lanes, SliceHost& h);

// Tiled schedule of the natural-order triangular sweeps.  Rows are grouped into units: the rows of one tile of the grid
// (boxes along the pattern's dominant lower offsets) that share a tile level, where tileLevel(i) is the maximum over
// the rows j < i coupled to i of tileLevel(j), plus one when j lies in another tile.  A unit therefore depends only on
// units of earlier tile levels and on its own rows, whatever the tiles are (the tiles only decide how many launches
// and local levels there are); inside a unit the rows are ordered by their local level.
struct TileSched {
    int dims[3] = {1, 1, 1};
    std::vector<int> ord;          // working order: by tile level, tile, local level, row
    std::vector<int> unitStart;    // first working row of each unit (units + 1 entries)
    std::vector<int> tlStart;      // first unit of each tile level (tile levels + 1 entries)
    std::vector<int> lvPtr;        // unit u's local level starts: lvStart[lvPtr[u] .. lvPtr[u + 1]), offsets from the
    std::vector<int> lvStart;      // unit's first row, the last one its row count
    int maxRows = 0, maxLevels = 0;
    int tileLevels() const { return int(tlStart.size()) - 1; }
    int units() const { return int(unitStart.size()) - 1; }
};
// up to three dominant lower offsets of the pattern (ascending; returns how many) and per row its chain coordinate
// along each: the number of consecutive steps back by that offset present in the pattern (for a natural-ordered grid
// the cell's grid indices)
int gridCoords(const HostBsr& A, int off[3], std::vector<int> coord[3]);
void tileOrder(const HostBsr& A, const std::vector<int> coord[3], int ndim, const int dims[3], TileSched& T);

// Red / black tiles: the grid cut into tiles of dims cells along the pattern's dominant lower offsets, the tiles
// coloured by the parity of their tile coordinates; the bulk rows by colour, tile, local level (natural order inside a
// tile) and row, then the rows far longer than the bulk (a well's perforated cells after the Schur fill) in the level
// order of their own couplings.  A tile of one colour couples only to tiles of the other (face neighbours of a 7-point
// stencil), so the tiles of a colour are independent and each is a natural-order triangular sweep of its own.
struct RbTiles {
    std::vector<int> ord;          // working row -> original row
    int nRed = 0, nTile = 0;       // red bulk rows [0, nRed), black [nRed, nTile), far rows [nTile, n)
    std::vector<int> unitStart;    // (tile, local level) units of the bulk rows: first rows (+ nTile)
    std::vector<int> unitLevel;    // per unit: its factorisation level (red local levels, then black ones)
    std::vector<int> tileUnit;     // per tile (working order): first unit (+ units); tiles [0, redTiles) are red
    int redTiles = 0, redLevels = 0, blackLevels = 0;
    std::vector<int> farStart;     // level starts of the far rows, relative to nTile
};
// false when the pattern does not admit it (no grid offsets, one colour empty, same-colour tiles coupled)
bool rbTileOrder(const HostBsr& S, const int dims[3], RbTiles& T);

// AMG set-up on the host (AMG::setup, amg.cpp:137-252): returns the hierarchy
struct AmgLevelHost {
    HostCsr A;
    std::vector<int> agg;         // fine -> aggregate (empty on the coarsest level)
    int nc = 0;
    // Galerkin maps to recompute the next level's values from this level's values on the device:
    //   ap[t] = sum_{u in [apPtr[t], apPtr[t+1])} A.v[apSrc[u]]           (entries of A P, per fine row)
    //   next.v[q] = sum_{u in [cPtr[q], cPtr[q+1])} ap[cSrc[u]]            (R (A P), fine rows ascending)
    std::vector<int> apPtr, apSrc, cPtr, cSrc;
};
void amgSetup(HostCsr A0, std::vector<AmgLevelHost>& H);

// ------------------------------------------------------------------ kernel launchers (rsh_kernels.hip)
void launchPackSlices(hipStream_t s, int nb, const double* src, const int* srcIdx, long long nslots, void* dst, bool f32);
void launchPackLong(hipStream_t s, int nb, const double* src, const int* srcIdx, int count, void* dst, bool f32);
void launchPackColumn0(hipStream_t s, int nb, const double* src, const int* srcIdx, long long nslots, double* dst);
void launchGatherBlocks(hipStream_t s, int nb, const double* src, const int* map, int count, double* dst);

// gate (batched systems): with ks set, a slice / long row whose owner system is no longer iterating is skipped
// y = A x over a slice layout (all levels at once); TA value type, TV vector type
void launchSpmvSliced(hipStream_t s, int nb, bool f32, const SliceDev<double>* d64, const SliceDev<float>* d32, const double* x,
                      double* y, const KScalars* gate = nullptr);
// precision of an ILU apply: FP64 factors and vectors; FP32 factors and FP32 working vector (FP32 arithmetic);
// FP32-stored factors with the FP64 working vector and FP64 arithmetic (an exactly linear FP64 operator)
enum IluFp { kIluFp64 = 0, kIluFp32 = 1, kIluStore32 = 2 };
// ILU sweeps of one level (fp: IluFp).  zFused (forward): the rows whose slice lane carries a pivot block (dsrc >= 0:
// no upper entries, read by no later row) finish their backward step here, z = D^-1 y into y and z (+ xp)
// red / black tile sweeps: tiles[t] = (first slice, end slice, first working row, end row) of tile t, one slice per
// local level, at most kTileWaveCap slots per row; forward (fwd) or backward
constexpr int kTileWaveCap = 8;
void launchIluTileWave(hipStream_t s, int nb, int fp, const void* slices, bool fwd, int ntiles, const int4* tiles, const double* r, void* y,
                       double* z, const double* xp);
void launchIluForward(hipStream_t s, int nb, int fp, const void* slices, int s0, int s1, int q0, int q1, const double* r, void* y,
                      const KScalars* gate = nullptr, double* zFused = nullptr, const double* xp = nullptr);
void launchIluBackward(hipStream_t s, int nb, int fp, const void* slices, int s0, int s1, int q0, int q1, void* y, double* z,
                       const double* xp, const KScalars* gate = nullptr, bool writeY = true);
// a run of small levels [l0, l1) swept by one workgroup (level starts on the device); its rows are the working rows
// [wbase, wbase + wcount), wcount <= kRunRowsMax, perm maps working rows to original rows
constexpr int kRunRowsMax = 2048;
void launchIluFused(hipStream_t s, int nb, int fp, const void* slices, bool fwd, int l0, int l1, const int* sStart, const int* qStart,
                    int wbase, int wcount, const int* perm, const double* r, void* y, double* z, const double* xp);
// tiled natural-order sweeps (TileSched): one launch per tile level [u0, u0 + nu) of units, one workgroup of kTileWaves
// wavefronts per unit (one per slice of the unit, one for all of its long rows).  rowsCap / levelsCap: the largest
// unit's rows and local levels; longMax: long rows per unit; longCap: inside slots staged per long row
constexpr int kTileWaves = 12;
constexpr int kTileLdsMax = 64 * 1024;
int tileLdsBytes(int nb, int fp, int rowsCap, int levelsCap, int longMax, int longCap);
void launchIluTile(hipStream_t s, int nb, int fp, const void* slices, bool fwd, int u0, int nu, const int* unitStart, const int* sStart,
                   const int* qStart, const int* lvPtr, const int* lvStart, int rowsCap, int levelsCap, int longMax, int longCap, const int* perm,
                   const double* r, void* y, double* z, const double* xp);
// dense groups of a run: largest group (block rows); the dense product runs one workgroup of kDenseThreads per
// tile of kWave scalar rows of a group
constexpr int kDenseMaxRows = 256;
constexpr int kDenseThreads = 1024;
// a run swept through its dense groups: outside contributions of the tasks [t0, t0+ns) / long rows [u0, u0+nq) into
// rt, then per task {group, first scalar row} the dense product into y (and z after the backward sweep)
void launchRunDense(hipStream_t s, int nb, int fp, const void* slices, bool fwd, int t0, int ns, int u0, int nq, int wbase, int wend,
                    int ntasks, const int2* tasks, const int* start, const int* rows, const long long* off, const void* inv,
                    const int* perm, const double* r, void* rt, void* y, double* z, const double* xp, const int* gowner = nullptr,
                    const KScalars* gate = nullptr, void* ycopy = nullptr, bool inv64 = false);
// numeric set-up of the dense groups of one sweep direction: gather the factor blocks, invert per group
void launchDenseGroups(hipStream_t s, int nb, bool upper, int ngroups, int maxRows, const int* start, const long long* off, int nmap, const int4* map,
                       const double* lu, double* raw, long long rawElems, double* inv);
void launchToFloat(hipStream_t s, long long n, const double* a, float* b);
// well Schur reduction of a scalar Jacobian on the device (see rsh_wells.cpp for the maps)
void launchWellDinv(hipStream_t s, int nw, const int* wdiag, const int* wcPtr, const int* wcSrc, const int* wwPtr, const int* wwSrc,
                    const double* jval, double* dinv, int* bad);
void launchSchurValues(hipStream_t s, long long nvals, const int* direct, const int* termPtr, const int4* terms, const double* jval,
                       const double* dinv, double* sval);
void launchSchurRhs(hipStream_t s, int nc, const int* cwPtr, const int* cwWell, const int* cwSrc, const double* jval, const double* dinv,
                    const double* b, double* bc);
void launchSchurRecover(hipStream_t s, int nc, int nw, const int* wcPtr, const int* wcCol, const int* wcSrc, const double* jval, const double* dinv,
                        const double* b, const double* xc, double* x);
// block-Jacobi second stage: D_b^-1 per block row (diag positions of a natural BSR), and z = D_b^-1 t + E xp
void launchBlockDiagInv(hipStream_t s, int nb, int n, const int* diag, const double* val, double* out);
void launchBlockJacobi(hipStream_t s, int nb, int n, const void* dinv, bool f32, const double* t, const double* xp, double* z);
// BSR values with block row i scaled by (1 + amp * u_i), u_i uniform in [-1, 1) hashed from (seed, i)
void launchPerturbRows(hipStream_t s, int n, const int* rp, int nb, const double* base, double* out, double amp, unsigned long long seed);
void launchGsFused(hipStream_t s, int l0, int l1, const int* lstart, const int* rows, const int* rp, const int* ci, const int* diag, const double* v,
                   const double* b, const double* xin, double* xout, bool forward);
// BILU0 elimination of long rows rows[r0..r1): one wavefront per row.  With map (device arrays): the row held in LDS
// and its updates taken from the precomputed pairs (for long row t = rows index, lower entry p of row i:
// map[mptr[mrow[t] + p - rp[i]] ..] = {U block q of row ci[p], entry offset s in row i}); maxLen = the longest row
struct LongRowMap { const int* mrow; const int* mptr; const int2* map; int maxLen; };
constexpr int kLongRowLdsMax = 48 * 1024;
void launchBilu0LongRows(hipStream_t s, int nb, const int* rows, int r0, int r1, const int* rp, const int* ci, const int* diag, double* val, int* nreg,
                         bool lean = false, const LongRowMap* map = nullptr);
// BILU0 numeric factorisation of the rows of one level (in place, natural layout); stop >= 0: only the lower entries
// in columns below stop, and the pivot is left for kIluGroupFactor
// (lean: every finished row is stored rounded to float32, the C45 lean factorisation)
void launchBilu0Level(hipStream_t s, int nb, const int* rows, int r0, int r1, const int* rp, const int* ci, const int* diag,
                      double* val, int* nreg, int stop = -1, bool lean = false);
// right-looking elimination of the groups [g0, g1) (members grows[gstart[g] .. gstart[g+1]) in ascending row order,
// member-pair block table gidx at gidxOff[g]), one workgroup per group
void launchIluGroupFactor(hipStream_t s, int nb, int g0, int g1, const int* gstart, const int* grows, const long long* gidxOff, const int* gidx,
                          const int* diag, double* val, int* nreg, const int* rp, bool lean);
// x = double(float(x)) in place (the lean factorisation's rounded input)
void launchRoundToFloat(hipStream_t s, long long n, double* x);
// CPR kernels
// quasi-IMPES weights, or true-IMPES from n accumulation blocks acc (ill-conditioned cells counted in fallbacks)
void launchCprWeights(hipStream_t s, int nb, int n, const int* diag, const double* val, double* w, const double* acc = nullptr,
                      int* fallbacks = nullptr);
void launchCprPressure(hipStream_t s, int nb, int n, const int* rp, const double* val, const double* w, double* ap);
void launchFixDiag(hipStream_t s, int n, const int* diag, double* v);
// r_p = w^T r; with x1 given also level 0's first damped-Jacobi iterate x1 = omega D^-1 r_p (dinv FP32 when f32)
void launchCprRestrict(hipStream_t s, int nb, int n, const double* w, const double* r, double* rp, const void* dinv = nullptr,
                       bool f32 = false, double omega = 0, double* x1 = nullptr);
// colP: the pressure column in the SpMV slice layout, FP64 or (f32) FP32 storage; y0 (optional): also t of the rows
// with iperm[row] < w1 into y0[iperm[row]] (the ILU's first level when it has no lower entries)
void launchCprSecond(hipStream_t s, int nb, const SliceDev<double>* sl, const void* colP, bool f32, const double* r, const double* xp,
                     double* t, const int* iperm = nullptr, int w1 = 0, double* y0 = nullptr);
// scalar AMG level kernels
void launchCsrSpmvResidual(hipStream_t s, int n, const int* rp, const int* ci, const double* v, const double* x, const double* b,
                           double* r);
void launchGsLevel(hipStream_t s, int nrows, const int* rows, const int* rp, const int* ci, const int* diag, const double* v,
                   const double* b, const double* xin, double* xout, bool forward);
// meanAgg: mean aggregate size (sets the lanes per aggregate); with x1 given also the coarse level's first
// damped-Jacobi iterate x1 = omega D^-1 bc
void launchRestrict(hipStream_t s, int nc, const int* rrp, const int* rci, const double* r, double* bc, double meanAgg,
                    const void* dinv = nullptr, bool f32 = false, double omega = 0, double* x1 = nullptr);
// x1 = omega D^-1 b
void launchJacobiInit(hipStream_t s, int n, bool f32, const void* dinv, const double* b, double omega, double* x1);
// damped-Jacobi steps from the first iterate x1: mode 0 out0 = b - A x1 (residual); mode 1 out0 = x1 + w D^-1 (b - A x1).
// sl != null: level 0 through the SpMV slices (values sval / lval); else CSR (rp, ci, v) with meanRow entries per row
// f32: level values (sval / lval / v) and dinv stored in FP32 (vectors and sums FP64).  pfine (CSR levels): also
// pfine[j] += out0[a] for the members j of aggregate a (prp / pci), the prolongation into the finer level's iterate
// lanes per CSR row of the coarse-level kernels for a mean row length (a power of two from 4 to 64)
int groupLanes(double mean);
void launchJacobiFused(hipStream_t s, int mode, bool f32, const SliceDev<double>* sl, const void* sval, const void* lval, int n, const int* rp,
                       const int* ci, const void* v, const void* dinv, const double* b, const double* x1, double omega, double* out0,
                       double meanRow, const int* prp = nullptr, const int* pci = nullptr, double* pfine = nullptr);
void launchDinv(hipStream_t s, int n, const int* diag, const double* v, double* dinv);
void launchProlongAdd(hipStream_t s, int n, const int* agg, const double* xc, double* x);
void launchDenseGemv(hipStream_t s, int m, const double* M, const double* x, double* y);
// with pfine given, also pfine[j] += y[a] for the members j of aggregate a (prp / pci)
void launchDenseGemv2(hipStream_t s, int m, int ld, const double* M, const double* x, double* y, const int* prp = nullptr,
                      const int* pci = nullptr, double* pfine = nullptr);
void launchGalerkin(hipStream_t s, int nap, const int* apPtr, const int* apSrc, const double* fine, double* ap, int nnzc,
                    const int* cPtr, const int* cSrc, double* coarse);
void launchDenseFromCsr(hipStream_t s, int m, const int* rp, const int* ci, const double* v, double* M);
// steps 1: column k; steps 2: columns k and k + 1 in one launch (same values as two one-step launches)
void launchGaussJordanStep(hipStream_t s, int m, int k, int steps, const double* Min, double* Mout);
// vectors and reductions
void launchFill(hipStream_t s, double* x, double v, long long n);
void launchCopy(hipStream_t s, double* dst, const double* src, long long n);
void launchDots(hipStream_t s, Ctx& c, long long n, int k, const double* a0, const double* b0, const double* a1, const double* b1,
                double* out, int epilogue, KScalars* ks);
// the same with the partial sums in red (kRedBlocks * 2 doubles at least) instead of the context's buffer
void launchDots(hipStream_t s, double* red, long long n, int k, const double* a0, const double* b0, const double* a1, const double* b1,
                double* out, int epilogue, KScalars* ks);
void launchBicgS(hipStream_t s, double* red, long long n, KScalars* ks, const double* r, const double* v, double* sv);
void launchBicgXR(hipStream_t s, double* red, long long n, KScalars* ks, const double* ph, const double* sh, const double* sv,
                  const double* t, const double* rh, double* x, double* r);
void launchBicgP(hipStream_t s, long long n, const KScalars* ks, const double* r, const double* v, double* p);
void launchBicgS(hipStream_t s, Ctx& c, long long n, KScalars* ks, const double* r, const double* v, double* sv);
// x += alpha ph + omega sh, r = s - omega t, then ||r|| and the next iteration's rho = (rh, r) (epilogue kEpRnRho);
// after the early exit on ||s|| only x += alpha ph
void launchBicgXR(hipStream_t s, Ctx& c, long long n, KScalars* ks, const double* ph, const double* sh, const double* sv,
                  const double* t, const double* rh, double* x, double* r);
void launchAxpy(hipStream_t s, long long n, double a, const double* x, double* y);                      // y += a x
void launchAxpyDev(hipStream_t s, long long n, const double* a, double sign, const double* x, double* y); // y += sign*a[0] x
void launchScaleTo(hipStream_t s, long long n, const double* num, const double* x, double* y, bool invert);  // y = x / num[0]
void launchSub(hipStream_t s, long long n, const double* b, const double* ax, double* r);                // r = b - ax
// out[i] = (V_i, w) for i < k with V_i = V + i * ld; deterministic per-block partials then a fixed-order sum
void launchMultiDot(hipStream_t s, Ctx& c, long long n, int k, const double* V, long long ld, const double* w, double* out);
// w -= sum_i h[i] V_i  (h on the device)
void launchMultiAxpy(hipStream_t s, long long n, int k, const double* V, long long ld, const double* h, double* w);
// x += sum_i y[i] Z_i  (y on the device)
void launchMultiAxpyAdd(hipStream_t s, long long n, int k, const double* Z, long long ld, const double* y, double* x);
void launchCopyKernel(hipStream_t s, const double* src, double* dst, long long n);

// batched BiCGSTAB over independent systems: task = {system, first entry, end entry} of the flat vectors,
// tStart[s] = first task of system s; partial holds K * T sums
void launchBInit(hipStream_t s, int ns, const double* tol, KScalars* ks);
// per-system sums (into out[sys * K + k] when out is set) and the epilogue (when ks is set)
void launchBFinal(hipStream_t s, int ns, int K, const int* tStart, const double* partial, int T, int ep, KScalars* ks, double* out = nullptr);
void launchBAdd(hipStream_t s, int T, const int4* task, const double* d, double* x);   // x += d
void launchBMask(hipStream_t s, int T, const int4* task, const int* keep, double* v);  // v = 0 where keep[system] == 0
void launchBDots(hipStream_t s, int T, const int4* task, int K, const double* a0, const double* b0, const double* a1, const double* b1,
                 double* partial);
void launchBBicgP(hipStream_t s, int T, const int4* task, const KScalars* ks, const double* r, const double* v, double* p);
void launchBBicgS(hipStream_t s, int T, const int4* task, const KScalars* ks, const double* r, const double* v, double* sv, double* partial);
void launchBBicgXs(hipStream_t s, int T, const int4* task, const KScalars* ks, const double* ph, double* x);
void launchBBicgXR(hipStream_t s, int T, const int4* task, const KScalars* ks, const double* ph, const double* sh, const double* sv,
                   const double* t, double* x, double* r, double* partial);

}  // namespace rsh
