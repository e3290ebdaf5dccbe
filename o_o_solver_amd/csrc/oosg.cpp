// Host driver of the numeric multifrontal LDLᵀ on one GPU: the C ABI of include/oosg.h.
//
// The reference runs the supernodes of a subtree one after the other on a thread and the top of the tree in waves
// (MultifrontalDriver::run, src/linalg/supernodal/MultifrontalDriver.cpp:651-702). Here the tree is cut into levels
// by height (leaves first), and every front of a level is processed by the same few launches: the small fronts by one
// fused kernel (one workgroup per front), the large ones by an assembly kernel, then panel steps of PANEL_BLOCK
// columns (diagonal block, rows below it, MFMA trailing update) and one MFMA launch for their update matrices. All
// launches of a factorisation are known from the symbolic structure alone and are listed once, in the plan.

#include "kernels.h"
#include "selector.h"

#include "../../include/oosg.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{

thread_local std::string g_lastError;
// What the last failure names besides its message: the equations of a SingularityError, the shortfall of a
// MemoryBudgetError.
thread_local std::vector<int32_t> g_lastEquations;
thread_local oosg_memory_shortfall g_lastShortfall{};

struct Failure
{
    oosg_status status;
    std::string message;
    std::vector<int32_t> equations;
    oosg_memory_shortfall shortfall{};
};

[[noreturn]] void fail(oosg_status status, const std::string& message)
{
    throw Failure{status, message, {}, {}};
}

[[noreturn]] void failNaming(oosg_status status, const std::string& message, std::vector<int32_t> equations)
{
    throw Failure{status, message, std::move(equations), {}};
}

void check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        fail(OOSG_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F> oosg_status guarded(F&& body)
{
    try
    {
        body();
        return OOSG_OK;
    }
    catch (const Failure& f)
    {
        g_lastError = f.message;
        g_lastEquations = f.equations;
        g_lastShortfall = f.shortfall;
        return f.status;
    }
    catch (const std::bad_alloc&)
    {
        g_lastError = "host memory exhausted";
        g_lastEquations.clear();
        g_lastShortfall = {};
        return OOSG_OUT_OF_MEMORY;
    }
    catch (const std::exception& e)
    {
        g_lastError = e.what();
        g_lastEquations.clear();
        g_lastShortfall = {};
        return OOSG_INTERNAL;
    }
}

// Device memory owned by a handle. A buffer counted in a plan's budget holds the plan's count (`account`) and gives
// its bytes back when it is freed.
struct DeviceBuffer
{
    void* ptr = nullptr;
    size_t bytes = 0;
    uint64_t* account = nullptr;
    bool owned = true; // false: a view of memory held elsewhere (never freed here)
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes), account(o.account), owned(o.owned)
    {
        o.ptr = nullptr;
        o.bytes = 0;
        o.account = nullptr;
        o.owned = true;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        reset();
        ptr = o.ptr;
        bytes = o.bytes;
        account = o.account;
        owned = o.owned;
        o.ptr = nullptr;
        o.bytes = 0;
        o.account = nullptr;
        o.owned = true;
        return *this;
    }
    ~DeviceBuffer() { reset(); }
    // A view of `size` bytes at `at` (not owned, not counted).
    static DeviceBuffer view(void* at, size_t size)
    {
        DeviceBuffer b;
        b.ptr = at;
        b.bytes = size;
        b.owned = false;
        return b;
    }
    void reset()
    {
        if (ptr != nullptr && owned)
            (void)hipFree(ptr);
        if (account != nullptr)
            *account -= bytes;
        ptr = nullptr;
        bytes = 0;
        account = nullptr;
        owned = true;
    }
    void allocate(size_t size, const char* what)
    {
        reset();
        if (size == 0)
            size = 8;
        check(hipMalloc(&ptr, size), what);
        bytes = size;
    }
    template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

template <typename T> DeviceBuffer upload(const std::vector<T>& host, hipStream_t stream, const char* what)
{
    DeviceBuffer b;
    b.allocate(host.size() * sizeof(T), what);
    if (!host.empty())
        check(hipMemcpyAsync(b.ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, stream), what);
    return b;
}

enum class LaunchKind
{
    Sync, // no kernel: only the launch's wait and record
    Small,
    Assemble,
    Diagonal,
    Tiles,
    Inverse // a round of the wide supernodes' triangle inverses
};

// The kernels as the profile counts them.
enum ProfileKind
{
    PROFILE_SMALL = 0,
    PROFILE_ASSEMBLE,
    PROFILE_DIAGONAL,
    PROFILE_PANEL_SOLVE,
    PROFILE_PANEL_UPDATE,
    PROFILE_UPDATE_MATRIX,
    PROFILE_FORWARD,           // forward sweep: gathers and the triangles of supernodes not wide
    PROFILE_BACKWARD,          // backward sweep: the triangles of supernodes not wide
    PROFILE_RESIDUAL,          // residual products
    PROFILE_FORWARD_ROWS,      // forward sweep: the L_RF row tiles
    PROFILE_BACKWARD_COLUMNS,  // backward sweep: the L_RFᵀ column blocks
    PROFILE_TRIANGLE_PRODUCTS, // both sweeps: the wide triangles, products with their whole inverses
    PROFILE_TRIANGLE_INVERSES, // factorisation: the whole inverses of the wide triangles
    PROFILE_UPDATE_MATRIX_BIG, // factorisation: update matrices on 128 × 128 tiles
    PROFILE_PANEL_UPDATE_FUSED, // factorisation: panel updates that also factorise the next diagonal blocks
    PROFILE_KINDS
};
static_assert(PROFILE_KINDS == OOSG_PROFILE_KINDS, "the profile kinds of the header");
static_assert(oosg::SCRATCH_BYTES == OOSG_NOALLOC_SCRATCH_BYTES, "the scratch array of the header");

struct Launch
{
    LaunchKind kind;
    int profile = PROFILE_SMALL;
    int64_t offset = 0; // into the list of its kind
    int count = 0;
    int k0 = 0;
    int maxOrder = 0;
    double flops = 0.0; // algorithmic flops of the launch
    // Assembly launches: the values they write (the panels' lower parts) and read (the children's update-matrix parts
    // landing there), the algorithmic bytes per value size; the matrix's entries they read are not counted.
    double values = 0.0;
    // Update-matrix launches: the same products as 128 × 128 tiles (the big tile list), and whether each precision
    // runs them so (the tile policy and the autotuner set it).
    int64_t bigOffset = 0;
    int bigCount = 0;
    bool big[2] = {false, false};
    int level = 0;
    // The stream it runs on (LANE_CRITICAL, LANE_BULK, LANE_INVERSES), and the events it waits for before and
    // records after it on that stream (-1: none).
    int lane = 0, wait = -1, record = -1;
    bool fuses = false; // tile launches: some tiles factorise the diagonal block they finish (mode 2)
};

// The factorisation's streams: the critical path (diagonal blocks, panel solves, the next panel's update, assembly
// and update matrices) on the plan's stream; the rest of each step's trailing update beside it; the whole inverses of
// a level's wide triangles beside both, each of their steps once the panel step it reads is done.
enum
{
    LANE_CRITICAL = 0,
    LANE_BULK = 1,
    LANE_INVERSES = 2,
    LANES = 3
};
enum
{
    EVENT_SOLVED = 0, // a step's panel solve is done: the bulk update may start
    EVENT_RESTED,     // the bulk update of a step is done
    EVENT_PANELS,     // a level's panels are final: the last launch of its whole inverses may start
    EVENT_FORK,
    EVENT_JOIN_BULK,
    EVENT_JOIN_INVERSES,
    EVENTS
};

// Per-launch device times by HIP events on the plan's stream, summed by kernel kind.
struct Profiler
{
    bool enabled = false;
    std::vector<hipEvent_t> events;
    std::vector<std::pair<int, double>> pending; // (kind, work) of each recorded launch
    double seconds[PROFILE_KINDS] = {};
    double work[PROFILE_KINDS] = {}; // algorithmic flops (factor kernels) or bytes (sweeps, residuals)
    int64_t launches[PROFILE_KINDS] = {};

    ~Profiler()
    {
        for (hipEvent_t e : events)
            (void)hipEventDestroy(e);
    }
    void reset()
    {
        pending.clear();
        for (int k = 0; k < PROFILE_KINDS; ++k)
        {
            seconds[k] = work[k] = 0.0;
            launches[k] = 0;
        }
    }
    // Records the event before a launch; returns its slot (or -1 when disabled).
    int begin(int kind, double launchWork, hipStream_t stream)
    {
        if (!enabled)
            return -1;
        const size_t slot = pending.size();
        while (events.size() < 2 * (slot + 1))
        {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess)
                return -1;
            events.push_back(e);
        }
        pending.emplace_back(kind, launchWork);
        (void)hipEventRecord(events[2 * slot], stream);
        return static_cast<int>(slot);
    }
    void end(int slot, hipStream_t stream)
    {
        if (slot >= 0)
            (void)hipEventRecord(events[2 * slot + 1], stream);
    }
    // After the stream is synchronised.
    void collect()
    {
        for (size_t k = 0; k < pending.size(); ++k)
        {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, events[2 * k], events[2 * k + 1]) != hipSuccess)
                continue;
            const int kind = pending[k].first;
            seconds[kind] += ms * 1e-3;
            work[kind] += pending[k].second;
            ++launches[kind];
        }
        pending.clear();
    }
};

// A launch sequence replayed as a HIP graph: captured from the stream the first time, and again whenever the buffers
// it names (its key) change; otherwise one hipGraphLaunch replaces its launches.
struct GraphCache
{
    hipGraphExec_t exec = nullptr;
    std::vector<const void*> key;

    ~GraphCache() { clear(); }
    void clear()
    {
        if (exec != nullptr)
            (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
    template <typename F> void run(hipStream_t stream, std::vector<const void*> newKey, F&& body)
    {
        if (exec == nullptr || newKey != key)
        {
            clear();
            check(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "graph: capture");
            try
            {
                body();
            }
            catch (...)
            {
                hipGraph_t partial = nullptr;
                (void)hipStreamEndCapture(stream, &partial);
                if (partial != nullptr)
                    (void)hipGraphDestroy(partial);
                throw;
            }
            hipGraph_t graph = nullptr;
            check(hipStreamEndCapture(stream, &graph), "graph: capture");
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            check(e, "graph: instantiate");
            key = std::move(newKey);
        }
        check(hipGraphLaunch(exec, stream), "graph: launch");
    }
};

double frontFlops(int64_t c, int64_t r)
{
    double f = 0.0;
    for (int64_t m = r; m < r + c; ++m)
        f += static_cast<double>(m) * static_cast<double>(m + 1) + 2.0 * static_cast<double>(m);
    return f;
}

} // namespace

// The device memory of one factor: panels, inverses of the diagonal blocks, pivots, and (FP32) the equations' scales.
// Sized for FP64 values, so that either precision can use it; a mixed factor's FP32 panels and inverses besides, once
// one was made (kept with the storage for the next mixed factor).
struct FactorStorage
{
    DeviceBuffer panels, linv, diag, scale, panels32, linv32;
};

struct oosg_plan
{
    int device = 0;
    // The device memory budget and what the plan and its live factors hold of it (declared first: the buffers below
    // give their bytes back to it when they are freed).
    uint64_t budget = 0;
    uint64_t reserved = 0;
    // The caller's stream (calls that read the caller's device memory wait for its work first) and its event.
    hipStream_t callerStream = nullptr;
    hipEvent_t callerEvent = nullptr;
    // Cancellation: polled between the parts of a call (null: none).
    oosg_cancel_fn cancelled = nullptr;
    void* cancelContext = nullptr;
    // This device's copy of the no-allocation solve's scratch array (oosg::scratchAddress), taken at plan creation.
    void* scratch = nullptr;
    // Its holders: the caller that made it, and every matrix and factor made from it (the last release deletes it, so
    // that a matrix or factor released after the plan still finds it).
    int32_t references = 1;
    hipStream_t stream = nullptr;
    bool ownsStream = true; // false: the stream of another plan (the structure of a pivoted factor)
    hipStream_t lanes[LANES] = {nullptr, nullptr, nullptr}; // lanes[0] = stream
    hipEvent_t laneEvents[EVENTS] = {};
    // Whether the factorisation uses the bulk and inverse streams beside the critical one (set by the autotuner when
    // it measures faster; never while the profiler times single launches).
    bool concurrent[2] = {false, false}; // per precision
    // Whether the factorisation is launched one launch at a time even where graphs are on (per precision; set by the
    // autotuner where three streams launched so measure faster than any captured graph).
    bool eager[2] = {false, false};
    hipEvent_t start = nullptr, stop = nullptr;
    int32_t n = 0, S = 0;
    // A partial plan (PartialFactor): its last supernode, the trailing set (trailingDofs columns, no rows below), is
    // assembled into K* and never eliminated nor swept (-1: a whole plan); eliminated: the equations factorised. The
    // trailing front's assembly blocks, and per trailing column the children's update rows that land on it (child
    // order) for the condensation of loads.
    int32_t trailing = -1, trailingDofs = 0, eliminated = 0;
    // An out-of-core plan (OutOfCoreFactor): each level's panels and inverses are made in one of two device windows
    // (level l in window l mod 2; panelOff and linvOff are offsets within the level's window), spooled to pinned host
    // memory once the level is done (level l at spoolPanelBase[l], spoolLinvBase[l]) and streamed back level by level
    // by the solves. The copies run on copyStream; oocReady[w] is recorded once window w holds what its next user
    // needs, oocFree[w] once its last user is done with it.
    bool outOfCore = false;
    int64_t windowPanels = 0, windowLinv = 0;
    std::vector<int64_t> spoolPanelBase, spoolLinvBase;
    hipStream_t copyStream = nullptr;
    hipEvent_t oocReady[2] = {nullptr, nullptr}, oocFree[2] = {nullptr, nullptr};
    std::vector<oosg::AssemblyBlock> trailingBlocks;
    std::vector<oosg::AssemblyItem> trailingItems;
    std::vector<int32_t> trailingPullStart, trailingPulls;
    std::vector<int32_t> first, ncol, nrow, parent, rows, childStart, children, rel, levelStart, levelList;
    std::vector<int32_t> levelOf; // the level (height) of each supernode (-1: the trailing one)
    // Each level's supernodes split for the sweeps: the tiny ones (a wave each) and the others; the tiny ones' pull
    // table (oosg::DevicePlan::pulls) and its width.
    std::vector<int32_t> tinyList, tinyStart, otherList, otherStart, pullBase, pulls, mediumList, mediumStart;
    // The forward sweep's gathers of the other supernodes: their positions in otherList order (otherPosStart, one
    // more value), and per position the children's update rows that land there (gatherStart, gatherRows; child
    // order).
    std::vector<int64_t> otherPosStart; // per supernode (those not swept as tiny: -1)
    std::vector<int32_t> gatherStart, gatherRows, gatherChunkStart;
    std::vector<oosg::GatherChunk> gatherChunks; // by level: chunks of GATHER_CHUNK positions
    int32_t pullWidth = 0;
    std::vector<int64_t> rowStart, panelOff, updateOff, linvOff;
    // The sweeps' work on big supernodes, by level: row tiles (forward) and column blocks (backward), in chunks where
    // their products are long, with the subtractions of the chunks' partial sums; the most slots of partial sums a
    // level uses.
    std::vector<int4> forwardTiles, backwardBlocks, forwardSubtractions, backwardSubtractions;
    std::vector<int64_t> forwardTileStart, backwardBlockStart, forwardSubtractionStart, backwardSubtractionStart;
    int64_t partialSlots = 0;
    // The wide supernodes' triangle blocks (s, I), by level (the sweeps' products with the whole inverse), as chunk
    // items for each sweep {s, I, chunk, slot} with the assignments of their sums to X; and per level the rounds of
    // the factorisation that build the inverses.
    std::vector<int4> forwardTriangles, backwardTriangles, forwardTriangleSums, backwardTriangleSums;
    std::vector<int64_t> forwardTriangleStart, backwardTriangleStart, triangleSumStart;
    struct InverseRound
    {
        int step; // −1: the last launch
        int64_t offset;
        int count;
        double flops;
        // Before step K: row block K's final values (items (s, K, J < K), launched as step −1).
        int64_t rowOffset = 0;
        int rowCount = 0;
        double rowFlops = 0.0;
    };
    std::vector<int2> inverseItems;
    std::vector<std::vector<InverseRound>> inverseRounds;
    int32_t levels = 0;
    int64_t smallFronts = 0, largeFronts = 0;
    double flops = 0.0;

    std::vector<Launch> schedule;
    std::vector<int32_t> smallList, stepFronts;
    std::vector<oosg::AssemblyBlock> blocks;
    std::vector<oosg::AssemblyItem> items;
    std::vector<oosg::UpdateTile> tiles, bigTiles;
    // The update-matrix tiles' extend-add items (children's parts), by quarter: itemStarts[tile.items + q].
    std::vector<int32_t> itemStarts;
    std::vector<oosg::AssemblyItem> extendItems;
    // The choice of tile size for the update matrices: 0 64 × 64 everywhere, 1 128 × 128 everywhere, 2 by level
    // (measured by oosg_plan_autotune, or a rule of thumb before).
    int32_t tilePolicy = 2;
    bool tuned[2] = {false, false};
    // Per-launch device times of a tuning run (ms), when timeLaunches is set.
    bool timeLaunches = false;
    std::vector<float> launchMs;
    std::vector<hipEvent_t> launchEvents;

    DeviceBuffer dFirst, dNcol, dNrow, dPanelOff, dUpdateOff, dRowStart, dRows, dRel, dChildStart, dChildren,
        dLinvOff;
    DeviceBuffer dTinyList, dOtherList, dMediumList, dPullBase, dPulls, dOtherPosStart, dGatherStart, dGatherRows,
        dGatherChunks;
    DeviceBuffer dTrailingBlocks, dTrailingItems, dTrailingPullStart, dTrailingPulls;
    // A mixed factor's boundary update matrices staged for their widening, and the ranges of that copy.
    DeviceBuffer widenTemp, widenRanges;
    DeviceBuffer dSmallList, dStepFronts, dBlocks, dItems, dTiles, dBigTiles, dItemStarts, dExtendItems, dLevelList, dForwardTiles, dBackwardBlocks,
        dForwardSubtractions, dBackwardSubtractions,
        dForwardTriangles, dBackwardTriangles, dForwardTriangleSums, dBackwardTriangleSums, dInverseItems;
    oosg::DevicePlan dp{};

    DeviceBuffer updates; // the update matrices of a factorisation (FP64 size; FP32 uses half)
    // Factor storage not in use: reserved with the plan, and given back by destroyed factors for the next
    // factorisation (the role of SupernodalFactor::releaseStorage and FactorOptions::storage, SupernodalFactor.hpp:91-96).
    std::vector<FactorStorage> storagePool;
    DeviceBuffer pivotStatus;   // per supernode the first zero pivot of a factorisation (INT32_MAX: none)
    DeviceBuffer pivotStats;    // the pivots' statistics of the last factorisation (oosg::PivotStatistics)
    // A combined factorisation and solve: from the first launch of level aheadLevel on (-1: off), the forward sweep of
    // the levels below it (aheadK right-hand sides, in solveX) runs on sweepStream beside the factorisation. The
    // levels from overlapLevel up hold one supernode each: the latency-bound top of the tree, whose panel steps leave
    // most of the GPU to the sweep (-1: no such levels).
    int32_t overlapLevel = -1, aheadLevel = -1, aheadK = 0;
    hipStream_t sweepStream = nullptr;
    hipEvent_t aheadEvents[LANES + 1] = {};
    // The indefinite factorisation's working storage, kept for the next one: the fronts of each level, W (the
    // updated columns of a level's fronts), the candidates' permutations, D and subdiagonals, the front descriptors
    // and states, and the launch lists.
    std::vector<DeviceBuffer> pivotFronts;
    DeviceBuffer pivotW, pivotPerm, pivotDiag, pivotSub, pivotDesc, pivotStates, pivotList, pivotTiles, pivotItems,
        pivotMaps, pivotRanges, pivotInverseItems;
    DeviceBuffer solveX, solveU, solveD, solveR, solveNorms;
    DeviceBuffer solvePartials; // the sweeps' slots of partial sums (partialSlots × 64 × 16)
    // The conjugate-gradient vectors (row-major n × k): b, x, r, p, q; the Gram products' partial sums, the Gram
    // products and the step's coefficient blocks.
    DeviceBuffer cgB, cgX, cgR, cgP, cgQ, cgPartial, cgDots;
    // The pivot growth of each supernode in the last (positive definite) factorisation of this plan, once there was
    // one (growthRecorded): the learned precision selector's history of the plan.
    DeviceBuffer supernodeGrowth;
    bool growthRecorded = false;
    // Launch sequences replayed as graphs (off while the profiler times single launches): the factorisation per
    // precision, and the sweeps.
    bool graphs = true;
    GraphCache factorGraph[2], sweepGraph;
    std::vector<double> levelPanelValues; // per level: Σ c·(c + r), the panel values a sweep of the level reads
    // The algorithmic work of each sweep launch of a level: the values of L (or of an inverse) it reads once, and the
    // rows of k right-hand-side values it reads or writes once (× k·8 bytes at run time).
    struct SweepWork
    {
        double triangleValues = 0, rowValues = 0, wideValues = 0;
        double forwardRows = 0, backwardRows = 0, rowTileRows = 0, columnRows = 0, wideRows = 0;
    };
    std::vector<SweepWork> sweepWork;
    Profiler profiler;

    ~oosg_plan()
    {
        for (hipEvent_t e : launchEvents)
            (void)hipEventDestroy(e);
        factorGraph[0].clear();
        factorGraph[1].clear();
        sweepGraph.clear();
        if (start)
            (void)hipEventDestroy(start);
        if (stop)
            (void)hipEventDestroy(stop);
        for (hipEvent_t e : laneEvents)
            if (e)
                (void)hipEventDestroy(e);
        for (int lane = 1; lane < LANES; ++lane)
            if (lanes[lane] && ownsStream)
                (void)hipStreamDestroy(lanes[lane]);
        for (hipEvent_t e : aheadEvents)
            if (e)
                (void)hipEventDestroy(e);
        if (sweepStream)
            (void)hipStreamDestroy(sweepStream);
        if (copyStream)
            (void)hipStreamDestroy(copyStream);
        for (hipEvent_t e : oocReady)
            if (e)
                (void)hipEventDestroy(e);
        for (hipEvent_t e : oocFree)
            if (e)
                (void)hipEventDestroy(e);
        if (callerEvent)
            (void)hipEventDestroy(callerEvent);
        if (stream && ownsStream)
            (void)hipStreamDestroy(stream);
    }
};

struct oosg_matrix
{
    oosg_plan* plan = nullptr;
    int32_t n = 0;
    int64_t nnz = 0;
    std::vector<int64_t> offsets;
    std::vector<int32_t> rows;
    std::vector<double> diagonal; // a_jj, 0 when not stored
    std::vector<double> hostValues; // the values as last given
    std::vector<int64_t> csrFromCsc; // per CSR entry, its CSC entry
    double normInf = 0.0;
    DeviceBuffer dOffsets, dRows, dValues, dPosition, dColumn, dScale, dCsrOffsets, dCsrCols, dCsrValues;
    // The CSR's row groups: runs of up to 8 consecutive rows with the same columns (a node's degrees of freedom);
    // group g is rows [csrGroups[g], csrGroups[g + 1]).
    std::vector<int32_t> csrGroups;
    DeviceBuffer dCsrGroups;
    DeviceBuffer dDiagonal; // a_jj on the device (the ratio test)
    std::vector<double> scale;
    // The CSR entries' CSC entries on the device (for values combined on the device), made on first use, and the
    // key of ‖A‖∞ they are summed into.
    DeviceBuffer dCsrMap, dNormKey;
    // Whether hostValues (and the selector's statistics) lag behind values set on the device.
    bool hostStale = false;
};

// A factor with Bunch–Kaufman pivoting: its structure after the interchanges and delays, as a plan of its own (for
// the sweeps and the inverses) whose equations are numbered in pivot order, supernode after supernode.
struct PivotedFactor
{
    std::unique_ptr<oosg_plan> structure;
    DeviceBuffer perm; // pivot position → equation
    DeviceBuffer sub;  // the subdiagonal of D by pivot position (0 outside 2 × 2 blocks)
    std::vector<int32_t> order, eliminated; // per supernode: its front's order N and its e
    std::vector<int64_t> rowsStart;         // per supernode: the start of its panel rows
    std::vector<int32_t> panelRows;         // the equations of each front's positions after the interchanges
    std::vector<int32_t> pivotOf;           // equation → pivot position
    std::vector<double> diag, sub_;         // D and its subdiagonal by pivot position
    int64_t delayed = 0, twoByTwo = 0;
};

struct oosg_factor
{
    oosg_plan* plan = nullptr;
    int32_t precision = OOSG_FP64;
    std::unique_ptr<PivotedFactor> pivoted; // indefinite mode
    DeviceBuffer panels;
    DeviceBuffer linv;  // the inverses of the diagonal blocks of PANEL_BLOCK columns (L_JJ⁻¹)
    DeviceBuffer dDiag; // D as the solves divide by it (FP32: of the scaled matrix)
    // A mixed factor (OOSG_MIXED): the levels below fp32Levels in FP32 (these buffers, panelOff / linvOff in floats),
    // the others in FP64 (panels, linv), all on the scaled matrix.
    DeviceBuffer panels32, linv32;
    int32_t fp32Levels = 0;
    DeviceBuffer scale; // FP32: the equations' scales of the factorised matrix
    // The pivots on the host (FP32: of the matrix, d_j / s_j²): filled on first use (pivoted factors: at once).
    mutable std::vector<double> diag;
    oosg_factor_info info{};
    // The first OOSG_RECORDED_ISSUE_COUNT pivots failing the ratio test, when the factor was kept despite them.
    std::vector<oosg_pivot_issue> issues;
    // Out of core: the panels, then the inverses, of every level in pinned host memory (the spool), and what the
    // factorisation wrote to it and the solves read back.
    void* spool = nullptr;
    size_t spoolBytes = 0;
    mutable oosg_out_of_core_traffic traffic{};
    ~oosg_factor()
    {
        if (spool != nullptr)
            (void)hipHostFree(spool);
        if (pivoted || plan == nullptr || panels.ptr == nullptr)
            return; // a pivoted factor is sized for its own structure
        FactorStorage back;
        back.panels = std::move(panels);
        back.linv = std::move(linv);
        back.diag = std::move(dDiag);
        back.scale = std::move(scale);
        back.panels32 = std::move(panels32);
        back.linv32 = std::move(linv32);
        if (plan->storagePool.size() < 2)
            plan->storagePool.push_back(std::move(back));
    }
};

namespace
{

// Whether a factor is of the diagonally scaled matrix and its solves refine by conjugate gradients (FP32 fronts, all
// of them or below a cut).
bool scaledPrecision(int precision)
{
    return precision != OOSG_FP64;
}

size_t valueBytes(int precision)
{
    return precision == OOSG_FP32 ? sizeof(float) : sizeof(double);
}

void setDevice(const oosg_plan* plan)
{
    check(hipSetDevice(plan->device), "hipSetDevice");
}

std::string mebibytes(uint64_t bytes)
{
    return std::to_string((bytes + (1u << 20) - 1) >> 20) + " MiB";
}

// MemoryBudgetError: `bytes` more do not fit the plan's budget (raised before the allocation, and before any numeric
// work of the call that asked).
[[noreturn]] void failBudget(const oosg_plan& p, uint64_t bytes, const std::string& what)
{
    Failure f{OOSG_OUT_OF_MEMORY,
              what + " needs " + mebibytes(bytes) + " of device memory; the budget of " + mebibytes(p.budget) +
                  " has " + mebibytes(p.budget > p.reserved ? p.budget - p.reserved : 0) + " left",
              {},
              {}};
    f.shortfall.requested_bytes = bytes;
    f.shortfall.available_bytes = p.budget > p.reserved ? p.budget - p.reserved : 0;
    f.shortfall.limit_bytes = p.budget;
    throw f;
}

// An allocation counted in the plan's budget (the buffer's previous memory is given back first).
void allocateCounted(oosg_plan& p, DeviceBuffer& b, size_t bytes, const char* what)
{
    b.reset();
    if (bytes == 0)
        bytes = 8;
    if (p.reserved + bytes > p.budget)
        failBudget(p, bytes, what);
    b.allocate(bytes, what);
    b.account = &p.reserved;
    p.reserved += bytes;
}

void ensureCounted(oosg_plan& p, DeviceBuffer& b, size_t bytes, const char* what)
{
    if (b.bytes < bytes)
        allocateCounted(p, b, bytes, what);
}

// The bytes of one factor's storage (FactorStorage), sized for FP64 values.
uint64_t factorStorageBytes(const oosg_plan& p);

// Storage for a factor: one given back to the plan, or a new allocation counted in the budget.
FactorStorage takeStorage(oosg_plan& p)
{
    const uint64_t panelBytes = static_cast<uint64_t>(std::max<int64_t>(p.panelOff[p.S], 1)) * sizeof(double);
    const uint64_t linvBytes = static_cast<uint64_t>(std::max<int64_t>(p.linvOff[p.S], 1)) * sizeof(double);
    const uint64_t vectorBytes = static_cast<uint64_t>(std::max(p.n, 1)) * sizeof(double);
    for (size_t k = 0; k < p.storagePool.size(); ++k)
    {
        FactorStorage& st = p.storagePool[k];
        if (st.panels.bytes >= panelBytes && st.linv.bytes >= linvBytes && st.diag.bytes >= vectorBytes &&
            st.scale.bytes >= vectorBytes)
        {
            FactorStorage out = std::move(st);
            p.storagePool.erase(p.storagePool.begin() + static_cast<std::ptrdiff_t>(k));
            return out;
        }
    }
    if (p.reserved + factorStorageBytes(p) > p.budget)
        failBudget(p, factorStorageBytes(p), "a factor (panels, inverses, pivots)");
    FactorStorage out;
    allocateCounted(p, out.panels, panelBytes, "factor: panels");
    allocateCounted(p, out.linv, linvBytes, "factor: inverses of the diagonal blocks");
    allocateCounted(p, out.diag, vectorBytes, "factor: pivots");
    allocateCounted(p, out.scale, vectorBytes, "factor: scales");
    return out;
}

// The plan's stream waits for the work queued on the caller's stream (before a call reads the caller's memory).
void followCaller(oosg_plan& p)
{
    if (p.callerEvent == nullptr)
        check(hipEventCreateWithFlags(&p.callerEvent, hipEventDisableTiming), "hipEventCreate");
    check(hipEventRecord(p.callerEvent, p.callerStream), "caller stream");
    check(hipStreamWaitEvent(p.stream, p.callerEvent, 0), "caller stream");
}

// OperationCancelledError (include/oos/base/errors/OperationCancelledError.hpp): the work queued is finished first.
void pollCancellation(oosg_plan& p)
{
    if (p.cancelled == nullptr || p.cancelled(p.cancelContext) == 0)
        return;
    for (hipStream_t s : p.lanes)
        if (s != nullptr)
            (void)hipStreamSynchronize(s);
    if (p.sweepStream != nullptr)
        (void)hipStreamSynchronize(p.sweepStream);
    fail(OOSG_CANCELLED, "the operation was cancelled (by an interruption or a progress callback)");
}

// ---- plan --------------------------------------------------------------------------------------------------------

void deriveLevels(oosg_plan& p);
oosg_status createPlan(const oosg_symbolic_desc* symbolic, int32_t trailing_dof_count, bool outOfCore, int32_t device,
                       uint64_t hbm_budget_bytes, oosg_plan** out);

void buildStructure(oosg_plan& p, const oosg_symbolic_desc& s)
{
    if (s.n < 0 || s.supernode_count < 0 || (s.supernode_count > 0 && (!s.first_column || !s.column_count ||
                                                                        !s.parent || !s.row_start)))
        fail(OOSG_INVALID_ARGUMENT, "plan: the symbolic factor is incomplete");
    p.n = s.n;
    p.S = s.supernode_count;
    const int32_t S = p.S;
    p.first.assign(s.first_column, s.first_column + S);
    p.ncol.assign(s.column_count, s.column_count + S);
    p.parent.assign(s.parent, s.parent + S);
    p.rowStart.assign(s.row_start, s.row_start + S + 1);
    if (p.rowStart[0] != 0)
        fail(OOSG_INVALID_ARGUMENT, "plan: row_start[0] must be 0");
    p.rows.assign(s.rows, s.rows + p.rowStart[S]);
    p.nrow.resize(S);
    int32_t expected = 0;
    for (int32_t k = 0; k < S; ++k)
    {
        if (p.first[k] != expected || p.ncol[k] < 1)
            fail(OOSG_INVALID_ARGUMENT, "plan: the supernodes must cover the columns in order, each with >= 1 column");
        expected += p.ncol[k];
        if (p.parent[k] != -1 && (p.parent[k] <= k || p.parent[k] >= S))
            fail(OOSG_INVALID_ARGUMENT, "plan: a parent must follow its child (postorder)");
        const int64_t r = p.rowStart[k + 1] - p.rowStart[k];
        if (r < 0 || r > INT32_MAX)
            fail(OOSG_INVALID_ARGUMENT, "plan: row starts must increase");
        p.nrow[k] = static_cast<int32_t>(r);
        int32_t last = p.first[k] + p.ncol[k] - 1;
        for (int64_t q = p.rowStart[k]; q < p.rowStart[k + 1]; ++q)
        {
            if (p.rows[q] <= last || p.rows[q] >= s.n)
                fail(OOSG_INVALID_ARGUMENT, "plan: the rows of a supernode must ascend below its columns");
            last = p.rows[q];
        }
        if (p.parent[k] == -1 && r > 0)
            fail(OOSG_INVALID_ARGUMENT, "plan: a root supernode has rows below it");
    }
    if (expected != s.n)
        fail(OOSG_INVALID_ARGUMENT, "plan: the supernodes do not cover the n columns");
    p.eliminated = p.n;
    if (p.trailingDofs > 0)
    {
        const int32_t T = S - 1;
        if (S < 1 || p.ncol[T] != p.trailingDofs || p.nrow[T] != 0)
            fail(OOSG_INVALID_ARGUMENT,
                 "plan: the trailing set of " + std::to_string(p.trailingDofs) +
                     " equations is not the last supernode (with no rows below it); analyse the symbolic factor "
                     "with the trailing nodes last");
        p.trailing = T;
        p.eliminated = p.n - p.trailingDofs;
    }
    // Children in increasing order.
    p.childStart.assign(S + 1, 0);
    for (int32_t k = 0; k < S; ++k)
        if (p.parent[k] >= 0)
            ++p.childStart[p.parent[k] + 1];
    for (int32_t k = 0; k < S; ++k)
        p.childStart[k + 1] += p.childStart[k];
    p.children.resize(p.childStart[S]);
    {
        std::vector<int32_t> fill(S, 0);
        for (int32_t k = 0; k < S; ++k)
            if (p.parent[k] >= 0)
                p.children[p.childStart[p.parent[k]] + fill[p.parent[k]]++] = k;
    }
    // Relative indices: each row of a child mapped into its parent's front (merge of two ascending lists).
    p.rel.assign(p.rows.size(), 0);
    for (int32_t k = 0; k < S; ++k)
    {
        const int32_t q = p.parent[k];
        if (q < 0)
            continue;
        const int32_t f = p.first[q], c = p.ncol[q];
        int64_t cursor = p.rowStart[q];
        for (int64_t e = p.rowStart[k]; e < p.rowStart[k + 1]; ++e)
        {
            const int32_t row = p.rows[e];
            if (row >= f && row < f + c)
                p.rel[e] = row - f;
            else
            {
                while (cursor < p.rowStart[q + 1] && p.rows[cursor] < row)
                    ++cursor;
                if (cursor >= p.rowStart[q + 1] || p.rows[cursor] != row)
                    fail(OOSG_INVALID_ARGUMENT, "plan: the rows of a child are not among its parent's columns and rows");
                p.rel[e] = c + static_cast<int32_t>(cursor - p.rowStart[q]);
            }
        }
    }
    deriveLevels(p);
}

// Splits each level for the sweeps into its tiny supernodes (isTinySupernode, and no front position hit by more than
// TINY_PULLS children's update rows) and the others, and builds the tiny ones' pull table: for each position of the
// front of a tiny supernode with children, the update rows that land there, in child order (the order in which the
// assembly adds them), padded with -1 to TINY_PULLS entries.
void splitTinySupernodes(oosg_plan& p)
{
    p.tinyList.clear();
    p.otherList.clear();
    p.mediumList.clear();
    p.mediumStart.assign(p.levels + 1, 0);
    p.pulls.clear();
    p.pullBase.assign(p.S, -1);
    p.tinyStart.assign(p.levels + 1, 0);
    p.otherStart.assign(p.levels + 1, 0);
    // Update rows are indexed by int32 in the table.
    const bool indexable = p.rowStart[p.S] < std::numeric_limits<int32_t>::max();
    std::vector<int32_t> hits;
    std::vector<char> tiny(p.S, 0);
    int32_t width = 0;
    int64_t positions = 0;
    for (int32_t s = 0; s < p.S; ++s)
    {
        if (!indexable || s == p.trailing || !oosg::isTinySupernode(p.ncol[s], p.nrow[s]))
            continue;
        const int32_t order = p.ncol[s] + p.nrow[s];
        hits.assign(order, 0);
        int32_t most = 0;
        for (int32_t q = p.childStart[s]; q < p.childStart[s + 1]; ++q)
        {
            const int32_t ch = p.children[q];
            for (int64_t e = p.rowStart[ch]; e < p.rowStart[ch + 1]; ++e)
                most = std::max(most, ++hits[p.rel[e]]);
        }
        if (most > oosg::TINY_PULLS)
            continue;
        tiny[s] = 1;
        width = std::max(width, most);
        if (p.childStart[s + 1] > p.childStart[s])
        {
            p.pullBase[s] = static_cast<int32_t>(positions);
            positions += order;
        }
    }
    if (positions * oosg::TINY_PULLS >= std::numeric_limits<int32_t>::max())
    {
        // A table too large for int32 offsets: no supernode is swept as tiny.
        std::fill(tiny.begin(), tiny.end(), 0);
        std::fill(p.pullBase.begin(), p.pullBase.end(), -1);
        width = 0;
        positions = 0;
    }
    p.pullWidth = width;
    p.pulls.assign(static_cast<size_t>(positions) * oosg::TINY_PULLS, -1);
    for (int32_t s = 0; s < p.S; ++s)
    {
        if (p.pullBase[s] < 0)
            continue;
        const int32_t order = p.ncol[s] + p.nrow[s];
        hits.assign(order, 0);
        for (int32_t q = p.childStart[s]; q < p.childStart[s + 1]; ++q)
        {
            const int32_t ch = p.children[q];
            for (int64_t e = p.rowStart[ch]; e < p.rowStart[ch + 1]; ++e)
            {
                const int32_t at = p.rel[e];
                p.pulls[(static_cast<size_t>(p.pullBase[s]) + at) * oosg::TINY_PULLS + hits[at]++] =
                    static_cast<int32_t>(e);
            }
        }
    }
    for (int32_t l = 0; l < p.levels; ++l)
    {
        for (int32_t q = p.levelStart[l]; q < p.levelStart[l + 1]; ++q)
        {
            const int32_t s = p.levelList[q];
            (tiny[s] ? p.tinyList : oosg::isMediumSupernode(p.ncol[s]) ? p.mediumList : p.otherList).push_back(s);
        }
        p.tinyStart[l + 1] = static_cast<int32_t>(p.tinyList.size());
        p.otherStart[l + 1] = static_cast<int32_t>(p.otherList.size());
        p.mediumStart[l + 1] = static_cast<int32_t>(p.mediumList.size());
    }
    // The gather lists of the medium and other supernodes (positions level by level) and their chunks by level.
    p.otherPosStart.assign(p.S, -1);
    int64_t gathered = 0;
    p.gatherChunks.clear();
    p.gatherChunkStart.assign(p.levels + 1, 0);
    std::vector<int32_t> gatheredList;
    for (int32_t l = 0; l < p.levels; ++l)
    {
        std::vector<int32_t> level(p.mediumList.begin() + p.mediumStart[l], p.mediumList.begin() + p.mediumStart[l + 1]);
        level.insert(level.end(), p.otherList.begin() + p.otherStart[l], p.otherList.begin() + p.otherStart[l + 1]);
        for (const int32_t s : level)
        {
            const int32_t order = p.ncol[s] + p.nrow[s];
            p.otherPosStart[s] = gathered;
            gathered += order;
            gatheredList.push_back(s);
            for (int32_t q0 = 0; q0 < order; q0 += oosg::GATHER_CHUNK)
                p.gatherChunks.push_back(oosg::GatherChunk{q0, order, p.ncol[s], p.first[s],
                                                           static_cast<int32_t>(p.rowStart[s]),
                                                           static_cast<int32_t>(p.otherPosStart[s]), 0, 0});
        }
        p.gatherChunkStart[l + 1] = static_cast<int32_t>(p.gatherChunks.size());
    }
    if (gathered >= std::numeric_limits<int32_t>::max() || p.rowStart[p.S] >= std::numeric_limits<int32_t>::max())
        fail(OOSG_INVALID_ARGUMENT, "plan: more front positions than the sweeps' gather lists hold");
    std::vector<int32_t> count(static_cast<size_t>(gathered) + 1, 0);
    for (const int32_t s : gatheredList)
    {
        for (int32_t q = p.childStart[s]; q < p.childStart[s + 1]; ++q)
        {
            const int32_t ch = p.children[q];
            for (int64_t e = p.rowStart[ch]; e < p.rowStart[ch + 1]; ++e)
                ++count[static_cast<size_t>(p.otherPosStart[s] + p.rel[e]) + 1];
        }
    }
    for (size_t q = 1; q < count.size(); ++q)
        count[q] += count[q - 1];
    p.gatherStart = count;
    p.gatherRows.assign(static_cast<size_t>(count.back()), 0);
    for (const int32_t s : gatheredList)
    {
        for (int32_t q = p.childStart[s]; q < p.childStart[s + 1]; ++q)
        {
            const int32_t ch = p.children[q];
            for (int64_t e = p.rowStart[ch]; e < p.rowStart[ch + 1]; ++e)
                p.gatherRows[count[static_cast<size_t>(p.otherPosStart[s] + p.rel[e])]++] = static_cast<int32_t>(e);
        }
    }
}

// The update matrices' places in one arena, by liveness: U(s) is written at the level of s and read at its parent's
// level, and its place is free again after that level (the role of the reference's stack of update matrices and the
// bound of its workspace, MultifrontalDriver.cpp:257-284, FactorSchedule.cpp:156; here a whole level is live at
// once, so the arena holds the update matrices of the levels in flight). Places are taken first fit, lowest address
// first, a level's largest matrices first, and freed places are merged. Every launch that reads or writes an update matrix runs on the critical stream
// in level order, so a place is reused only after the launches that read its previous matrix. updateOff[S] is the
// arena's size (elements).
void placeUpdateMatrices(oosg_plan& p, const std::vector<int32_t>& height)
{
    constexpr int64_t ALIGN = 32; // elements: 256 bytes in FP64
    const int32_t S = p.S;
    std::map<int64_t, int64_t> freeSpace; // start → length
    int64_t top = 0;
    std::vector<std::vector<int32_t>> releaseAfter(static_cast<size_t>(std::max(p.levels, 1)));
    const auto sizeOf = [&](int32_t s)
    { return (static_cast<int64_t>(p.nrow[s]) * p.nrow[s] + ALIGN - 1) / ALIGN * ALIGN; };
    for (int32_t l = 0; l < p.levels; ++l)
    {
        // A level's matrices placed largest first (less fragmentation than in supernode order).
        std::vector<int32_t> order(p.levelList.begin() + p.levelStart[l], p.levelList.begin() + p.levelStart[l + 1]);
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return sizeOf(a) > sizeOf(b); });
        for (const int32_t s : order)
        {
            const int64_t size = sizeOf(s);
            if (size == 0)
            {
                p.updateOff[s] = 0;
                continue;
            }
            auto it = freeSpace.begin();
            while (it != freeSpace.end() && it->second < size)
                ++it;
            if (it != freeSpace.end())
            {
                p.updateOff[s] = it->first;
                const int64_t rest = it->second - size, start = it->first + size;
                freeSpace.erase(it);
                if (rest > 0)
                    freeSpace[start] = rest;
            }
            else
            {
                p.updateOff[s] = top;
                top += size;
            }
            // Released once the parent's level has assembled it; the trailing front is assembled after every level
            // (its children's matrices stay to the end).
            if (p.parent[s] >= 0 && p.parent[s] != p.trailing)
                releaseAfter[height[p.parent[s]]].push_back(s);
        }
        for (const int32_t s : releaseAfter[l])
        {
            int64_t start = p.updateOff[s], length = sizeOf(s);
            auto next = freeSpace.lower_bound(start);
            if (next != freeSpace.end() && next->first == start + length)
            {
                length += next->second;
                next = freeSpace.erase(next);
            }
            if (next != freeSpace.begin())
            {
                auto previous = std::prev(next);
                if (previous->first + previous->second == start)
                {
                    start = previous->first;
                    length += previous->second;
                    freeSpace.erase(previous);
                }
            }
            if (start + length == top)
                top = start; // the end of the arena is free again
            else
                freeSpace[start] = length;
        }
    }
    // The arena's size: the highest end of any place.
    int64_t end = 0;
    for (int32_t s = 0; s < S; ++s)
        end = std::max(end, p.updateOff[s] + sizeOf(s));
    p.updateOff[S] = end;
}

// An out-of-core plan's layout: each level's panels (and inverses) from the start of its window, in supernode order;
// the two windows one after the other (panelOff[S], linvOff[S]: both windows); the spool's copy of each level.
void windowLayout(oosg_plan& p, const std::vector<int32_t>& height)
{
    const int32_t S = p.S;
    std::vector<int64_t> panelValues(S), linvValues(S);
    for (int32_t k = 0; k < S; ++k)
    {
        panelValues[k] = p.panelOff[k + 1] - p.panelOff[k];
        linvValues[k] = p.linvOff[k + 1] - p.linvOff[k];
    }
    std::vector<int64_t> panelFill(p.levels, 0), linvFill(p.levels, 0);
    for (int32_t k = 0; k < S; ++k)
    {
        if (k == p.trailing)
            continue;
        const int32_t l = height[k];
        p.panelOff[k] = panelFill[l];
        p.linvOff[k] = linvFill[l];
        panelFill[l] += panelValues[k];
        linvFill[l] += linvValues[k];
    }
    p.windowPanels = 0;
    p.windowLinv = 0;
    p.spoolPanelBase.assign(p.levels + 1, 0);
    p.spoolLinvBase.assign(p.levels + 1, 0);
    for (int32_t l = 0; l < p.levels; ++l)
    {
        // Windows of a multiple of 64 values (the second starts on a 512-byte boundary in FP64).
        p.windowPanels = std::max(p.windowPanels, (panelFill[l] + 63) / 64 * 64);
        p.windowLinv = std::max(p.windowLinv, (linvFill[l] + 63) / 64 * 64);
        p.spoolPanelBase[l + 1] = p.spoolPanelBase[l] + panelFill[l];
        p.spoolLinvBase[l + 1] = p.spoolLinvBase[l] + linvFill[l];
    }
    p.panelOff[S] = 2 * p.windowPanels;
    p.linvOff[S] = 2 * p.windowLinv;
}

// From the structure (first, ncol, nrow, parent, children): the offsets of the panels, update matrices and inverses,
// the levels by height, the flops, and the sweeps' work lists by level.
void deriveLevels(oosg_plan& p)
{
    const int32_t S = p.S;
    p.panelOff.assign(S + 1, 0);
    p.updateOff.assign(S + 1, 0);
    p.linvOff.assign(S + 1, 0);
    std::vector<int32_t> height(S, 0);
    for (int32_t k = 0; k < S; ++k)
    {
        const int64_t c = p.ncol[k], r = p.nrow[k];
        p.panelOff[k + 1] = p.panelOff[k] + c * (c + r); // the trailing supernode's panel holds K*
        // The diagonal blocks' inverses, then (wide supernodes) the whole inverse of the triangle; none for the
        // trailing supernode (not eliminated).
        const int64_t inverses =
            k == p.trailing
                ? 0
                : oosg::blockInverseValues(c) + (oosg::keepsTriangleInverse(static_cast<int>(c)) ? c * c : 0);
        p.linvOff[k + 1] = p.linvOff[k] + inverses;
        if (k != p.trailing)
            for (int64_t m = r; m < r + c; ++m)
                p.flops += static_cast<double>(m) * static_cast<double>(m + 1) + 2.0 * static_cast<double>(m);
        if (p.parent[k] >= 0)
            height[p.parent[k]] = std::max(height[p.parent[k]], height[k] + 1);
    }
    // The levels hold the supernodes eliminated (the trailing one is in none).
    p.levels = 0;
    for (int32_t k = 0; k < S; ++k)
        if (k != p.trailing)
            p.levels = std::max(p.levels, height[k] + 1);
    p.levelStart.assign(p.levels + 1, 0);
    for (int32_t k = 0; k < S; ++k)
        if (k != p.trailing)
            ++p.levelStart[height[k] + 1];
    for (int32_t l = 0; l < p.levels; ++l)
        p.levelStart[l + 1] += p.levelStart[l];
    p.levelList.resize(p.levelStart[p.levels]);
    p.levelOf.assign(S, -1);
    p.levelPanelValues.assign(p.levels, 0.0);
    p.sweepWork.assign(p.levels, oosg_plan::SweepWork{});
    {
        std::vector<int32_t> fill(p.levels, 0);
        for (int32_t k = 0; k < S; ++k)
        {
            if (k == p.trailing)
                continue;
            p.levelOf[k] = height[k];
            p.levelList[p.levelStart[height[k]] + fill[height[k]]++] = k;
            p.levelPanelValues[height[k]] += static_cast<double>(p.panelOff[k + 1] - p.panelOff[k]);
            oosg_plan::SweepWork& w = p.sweepWork[height[k]];
            const double c = p.ncol[k], r = p.nrow[k];
            double inBlocks = 0; // values of the diagonal blocks' lower triangles
            for (int q0 = 0; q0 < p.ncol[k]; q0 += oosg::PANEL_BLOCK)
            {
                const double kb = std::min(oosg::PANEL_BLOCK, p.ncol[k] - q0);
                inBlocks += kb * (kb + 1) / 2;
            }
            if (oosg::keepsTriangleInverse(p.ncol[k]))
            {
                w.wideValues += c * (c + 1) / 2; // the whole inverse's lower triangle
                w.wideRows += 2 * c;             // x_F read, W written
            }
            else
                w.triangleValues += oosg::blockInverseValues(p.ncol[k]) + (c * (c + 1) / 2 - inBlocks);
            // Gathers: each child's update vector read once, its target rows read and written; X_F read and written,
            // the update vector zeroed.
            double gathered = 0;
            for (int32_t q = p.childStart[k]; q < p.childStart[k + 1]; ++q)
                gathered += 3.0 * p.nrow[p.children[q]];
            if (oosg::sweepRowsInline(p.ncol[k], p.nrow[k]))
            {
                // L_RF in the triangle's workgroup: its values and the update vector's rows (written once more,
                // after the zeroing) in the forward launch; X_R gathered in the backward one.
                w.triangleValues += r * c;
                w.forwardRows += gathered + 2 * c + 2 * r;
                w.backwardRows += 2 * c + r;
                continue;
            }
            w.forwardRows += gathered + 2 * c + r;
            w.backwardRows += 2 * c;
            w.rowValues += r * c;
            w.rowTileRows += r > 0 ? 2 * r + c : 0;  // update vector read and written, x_F read
            w.columnRows += r > 0 ? r + 2 * c : 0;   // X_R read, X_F read and written
        }
    }
    // The sweeps' row tiles and column blocks of the big supernodes, level by level.
    p.forwardTileStart.assign(p.levels + 1, 0);
    p.backwardBlockStart.assign(p.levels + 1, 0);
    p.forwardSubtractionStart.assign(p.levels + 1, 0);
    p.backwardSubtractionStart.assign(p.levels + 1, 0);
    p.forwardTriangleStart.assign(p.levels + 1, 0);
    p.backwardTriangleStart.assign(p.levels + 1, 0);
    p.triangleSumStart.assign(p.levels + 1, 0);
    p.partialSlots = 0;
    if (p.rowStart[p.S] > INT32_MAX)
        fail(OOSG_INVALID_ARGUMENT, "plan: more than 2^31 rows below the supernodes");
    for (int32_t l = 0; l < p.levels; ++l)
    {
        int maxBlocks = 0;
        int32_t forwardSlots = 0, backwardSlots = 0; // each sweep's slots of this level
        for (int32_t q = p.levelStart[l]; q < p.levelStart[l + 1]; ++q)
        {
            const int32_t s = p.levelList[q];
            const bool wide = p.ncol[s] >= oosg::WIDE_COLUMNS;
            if (wide)
                maxBlocks = std::max(maxBlocks, (p.ncol[s] + 63) / 64);
            // The L_RF products of every supernode with rows below but the small ones (those do them in the
            // workgroup of their triangle): row tiles (forward), column blocks (backward). A supernode of a pivoted
            // factor whose columns were all delayed has none (c = 0).
            if (oosg::sweepRowsInline(p.ncol[s], p.nrow[s]) || oosg::isMediumSupernode(p.ncol[s]))
                continue;
            const int32_t c = p.ncol[s], r = p.nrow[s];
            const int32_t columnBlocks = (c + 63) / 64, rowTiles = (r + 63) / 64;
            const int32_t columnChunks = (columnBlocks + oosg::SWEEP_CHUNK - 1) / oosg::SWEEP_CHUNK;
            const int32_t rowChunks = (rowTiles + oosg::SWEEP_CHUNK - 1) / oosg::SWEEP_CHUNK;
            if (c > 0)
                for (int32_t t = 0; t < rowTiles; ++t)
                {
                    if (columnChunks == 1)
                    {
                        p.forwardTiles.push_back(make_int4(s, t, -1, -1));
                        continue;
                    }
                    // The row tile's columns in chunks, then the subtraction of their sums from its update rows.
                    for (int32_t ch = 0; ch < columnChunks; ++ch)
                        p.forwardTiles.push_back(make_int4(s, t, ch, forwardSlots + ch));
                    p.forwardSubtractions.push_back(make_int4(static_cast<int32_t>(p.rowStart[s]) + 64 * t,
                                                              std::min(64, r - 64 * t), forwardSlots,
                                                              columnChunks | (1 << 30)));
                    forwardSlots += columnChunks;
                }
            if (r > 0)
                for (int32_t b = 0; b < columnBlocks; ++b)
                {
                    if (rowChunks == 1)
                    {
                        p.backwardBlocks.push_back(make_int4(s, b, -1, -1));
                        continue;
                    }
                    for (int32_t ch = 0; ch < rowChunks; ++ch)
                        p.backwardBlocks.push_back(make_int4(s, b, ch, backwardSlots + ch));
                    p.backwardSubtractions.push_back(make_int4(p.first[s] + 64 * b, std::min(64, c - 64 * b),
                                                               backwardSlots, rowChunks));
                    backwardSlots += rowChunks;
                }
        }
        p.partialSlots = std::max<int64_t>(p.partialSlots, std::max(forwardSlots, backwardSlots));
        p.forwardTileStart[l + 1] = static_cast<int64_t>(p.forwardTiles.size());
        p.backwardBlockStart[l + 1] = static_cast<int64_t>(p.backwardBlocks.size());
        p.forwardSubtractionStart[l + 1] = static_cast<int64_t>(p.forwardSubtractions.size());
        p.backwardSubtractionStart[l + 1] = static_cast<int64_t>(p.backwardSubtractions.size());
        // Wide supernodes: their triangle's blocks (s, I) for the sweeps' products with the whole inverse, and the
        // rounds d of the factorisation that build that inverse (items (s, I), I >= d, in round d).
        int32_t forwardTriangleSlots = 0, backwardTriangleSlots = 0;
        for (int32_t q = p.levelStart[l]; q < p.levelStart[l + 1]; ++q)
        {
            const int32_t s = p.levelList[q];
            if (!oosg::keepsTriangleInverse(p.ncol[s]))
                continue;
            const int32_t nb = (p.ncol[s] + 63) / 64;
            for (int32_t I = 0; I < nb; ++I)
            {
                const int32_t rows = std::min(64, p.ncol[s] - 64 * I);
                // Forward: blocks J <= I; backward: J >= I; in chunks of SWEEP_CHUNK blocks.
                const int32_t forwardChunks = (I + 1 + oosg::SWEEP_CHUNK - 1) / oosg::SWEEP_CHUNK;
                const int32_t backwardChunks = (nb - I + oosg::SWEEP_CHUNK - 1) / oosg::SWEEP_CHUNK;
                for (int32_t ch = 0; ch < forwardChunks; ++ch)
                    p.forwardTriangles.push_back(make_int4(s, I, ch, forwardTriangleSlots + ch));
                for (int32_t ch = 0; ch < backwardChunks; ++ch)
                    p.backwardTriangles.push_back(make_int4(s, I, ch, backwardTriangleSlots + ch));
                p.forwardTriangleSums.push_back(
                    make_int4(p.first[s] + 64 * I, rows, forwardTriangleSlots, forwardChunks | (1 << 29)));
                p.backwardTriangleSums.push_back(
                    make_int4(p.first[s] + 64 * I, rows, backwardTriangleSlots, backwardChunks | (1 << 29)));
                forwardTriangleSlots += forwardChunks;
                backwardTriangleSlots += backwardChunks;
            }
        }
        p.partialSlots = std::max<int64_t>(p.partialSlots, std::max(forwardTriangleSlots, backwardTriangleSlots));
        p.forwardTriangleStart[l + 1] = static_cast<int64_t>(p.forwardTriangles.size());
        p.backwardTriangleStart[l + 1] = static_cast<int64_t>(p.backwardTriangles.size());
        p.triangleSumStart[l + 1] = static_cast<int64_t>(p.forwardTriangleSums.size());
        std::vector<oosg_plan::InverseRound> rounds;
        // Steps K = 0 .. maxBlocks − 2 (items (s, I, J), I > K >= J), each after row block K's final values (items
        // (s, K, J < K), for 1 <= K <= blocks − 2), then the last launch (the diagonal blocks and the last row block,
        // step −1).
        constexpr double ITEM_FLOPS = 2.0 * 64.0 * 64.0 * 64.0;
        for (int K = 0; K <= maxBlocks - 1; ++K)
        {
            const bool last = K == maxBlocks - 1;
            oosg_plan::InverseRound round{last ? -1 : (K % 2 == 1 ? K | (1 << 16) : K), 0, 0, 0.0};
            round.rowOffset = static_cast<int64_t>(p.inverseItems.size());
            for (int32_t q = p.levelStart[l]; !last && K >= 1 && q < p.levelStart[l + 1]; ++q)
            {
                const int32_t s = p.levelList[q];
                if (oosg::keepsTriangleInverse(p.ncol[s]) && K <= (p.ncol[s] + 63) / 64 - 2)
                    for (int J = 0; J < K; ++J)
                        p.inverseItems.push_back(make_int2(s, (K << 16) | J));
            }
            round.rowCount = static_cast<int>(p.inverseItems.size() - round.rowOffset);
            round.rowFlops = round.rowCount * ITEM_FLOPS;
            round.offset = static_cast<int64_t>(p.inverseItems.size());
            for (int32_t q = p.levelStart[l]; q < p.levelStart[l + 1]; ++q)
            {
                const int32_t s = p.levelList[q];
                const int blocks = (p.ncol[s] + 63) / 64;
                if (!oosg::keepsTriangleInverse(p.ncol[s]))
                    continue;
                // Steps in pairs: an even step K does the items of row block K + 1 only, the odd step K + 1 every
                // item below with the terms of both steps (each T_IJ read and written once for two steps).
                for (int I = 0; I < blocks; ++I)
                    for (int J = 0; J <= I; ++J)
                        if (last ? (I == J || I == blocks - 1) : (K % 2 == 0 ? I == K + 1 && J <= K : I > K && J <= K))
                        {
                            p.inverseItems.push_back(make_int2(s, (I << 16) | J));
                            round.flops += ITEM_FLOPS * (!last && K % 2 == 1 && J < K ? 2.0 : 1.0);
                        }
            }
            round.count = static_cast<int>(p.inverseItems.size() - round.offset);
            rounds.push_back(round);
        }
        p.inverseRounds.push_back(std::move(rounds));
    }
    placeUpdateMatrices(p, height);
    splitTinySupernodes(p);
    if (p.outOfCore)
        windowLayout(p, height);
}

// The algorithmic flops of a tile: 2·k per element it updates (the lower part only on the diagonal).
double tileFlops(const oosg::UpdateTile& t)
{
    double elements = static_cast<double>(t.m) * t.n;
    if (t.lower)
        elements -= static_cast<double>(t.n) * (t.n - 1) / 2.0;
    return 2.0 * t.k * elements;
}

Launch syncLaunch(int lane, int wait, int record)
{
    Launch l{LaunchKind::Sync, PROFILE_SMALL};
    l.lane = lane;
    l.wait = wait;
    l.record = record;
    return l;
}

// A launch of one step of a level's whole inverses, on the inverses' stream after the event `wait`.
// The launch of a step's row block finals (before the step's own launch).
Launch inverseRowsLaunch(const oosg_plan::InverseRound& round, int wait)
{
    Launch rows{LaunchKind::Inverse, PROFILE_TRIANGLE_INVERSES};
    rows.offset = round.rowOffset;
    rows.count = round.rowCount;
    rows.k0 = -1;
    rows.flops = round.rowFlops;
    rows.lane = LANE_INVERSES;
    rows.wait = wait;
    return rows;
}

Launch inverseLaunch(const oosg_plan::InverseRound& round, int wait)
{
    Launch inverse{LaunchKind::Inverse, PROFILE_TRIANGLE_INVERSES};
    inverse.offset = round.offset;
    inverse.count = round.count;
    inverse.k0 = round.step;
    inverse.flops = round.flops;
    inverse.lane = LANE_INVERSES;
    inverse.wait = wait;
    return inverse;
}

// The assembly blocks of the panel part (columns < c) of the front of s, each with the parts of its children's update
// matrices that land in it (items appended to `items`, indexed from its start); the update matrices take their
// children's parts in the update-matrix tiles (extendTilesKernel).
void appendPanelBlocks(const oosg_plan& p, int32_t s, std::vector<oosg::AssemblyBlock>& blocks,
                       std::vector<oosg::AssemblyItem>& items)
{
    const auto addBlock = [&](int j0, int j1, int i0, int i1)
    {
        oosg::AssemblyBlock blk{s, j0, j1, i0, i1, static_cast<int32_t>(items.size()), 0, 0};
        for (int32_t q = p.childStart[s]; q < p.childStart[s + 1]; ++q)
        {
            const int32_t ch = p.children[q];
            const int32_t* rel = p.rel.data() + p.rowStart[ch];
            const int32_t rc = p.nrow[ch];
            const auto at = [&](int value)
            { return static_cast<int32_t>(std::lower_bound(rel, rel + rc, value) - rel); };
            const oosg::AssemblyItem it{ch, at(j0), at(j1), at(i0), at(i1), 0};
            if (it.q0 < it.q1 && std::max(it.q0, it.ia) < it.ib)
                items.push_back(it);
        }
        blk.itemEnd = static_cast<int32_t>(items.size());
        blocks.push_back(blk);
    };
    const int c = p.ncol[s], n = c + p.nrow[s];
    int j0 = 0;
    while (j0 < c)
    {
        // Rows from the block's first column down: the panel's values above the diagonal are never read.
        const int i0 = j0;
        const int height = n - i0;
        if (height <= oosg::ASSEMBLY_BLOCK_VALUES)
        {
            int j1 = std::min(n, j0 + std::max(1, oosg::ASSEMBLY_BLOCK_VALUES / height));
            if (j0 < c && j1 > c)
                j1 = c;
            addBlock(j0, j1, i0, n);
            j0 = j1;
        }
        else
        {
            for (int i = i0; i < n; i += oosg::ASSEMBLY_BLOCK_VALUES)
                addBlock(j0, j0 + 1, i, std::min(n, i + oosg::ASSEMBLY_BLOCK_VALUES));
            j0 += 1;
        }
    }
}

// A partial plan's trailing front: its assembly blocks (K* = the matrix entries of its columns + its children's
// update matrices, in child order, as a large front's panel is assembled), and per trailing column the children's
// update rows that land on it, in child order (the condensation of loads adds them so).
void buildTrailing(oosg_plan& p)
{
    p.trailingBlocks.clear();
    p.trailingItems.clear();
    p.trailingPullStart.clear();
    p.trailingPulls.clear();
    if (p.trailing < 0)
        return;
    const int32_t T = p.trailing;
    appendPanelBlocks(p, T, p.trailingBlocks, p.trailingItems);
    const int32_t c = p.ncol[T];
    std::vector<int32_t> count(static_cast<size_t>(c) + 1, 0);
    for (int32_t q = p.childStart[T]; q < p.childStart[T + 1]; ++q)
    {
        const int32_t ch = p.children[q];
        for (int64_t e = p.rowStart[ch]; e < p.rowStart[ch + 1]; ++e)
            ++count[p.rel[e] + 1];
    }
    for (int32_t j = 0; j < c; ++j)
        count[j + 1] += count[j];
    p.trailingPullStart = count;
    p.trailingPulls.assign(static_cast<size_t>(count[c]), 0);
    for (int32_t q = p.childStart[T]; q < p.childStart[T + 1]; ++q)
    {
        const int32_t ch = p.children[q];
        for (int64_t e = p.rowStart[ch]; e < p.rowStart[ch + 1]; ++e)
            p.trailingPulls[count[p.rel[e]]++] = static_cast<int32_t>(e);
    }
}

void buildSchedule(oosg_plan& p)
{
    using oosg::PANEL_BLOCK;
    // Columns of a super-block: the trailing update beyond it is one product with K = SUPER_BLOCK.
    constexpr int SUPER_BLOCK = 4 * PANEL_BLOCK;
    // Panel steps between two records of EVENT_SOLVED (when the bulk update needs none sooner).
    constexpr int SOLVED_RECORD_STEPS = 4;
    // Every event is recorded once on the critical stream first, so that each wait refers to a record of the same
    // sequence (a captured graph has no outside dependencies).
    p.schedule.push_back(syncLaunch(LANE_CRITICAL, -1, EVENT_SOLVED));
    p.schedule.push_back(syncLaunch(LANE_CRITICAL, -1, EVENT_RESTED));
    p.schedule.push_back(syncLaunch(LANE_CRITICAL, -1, EVENT_PANELS));
    // The top levels of at most OVERLAP_FRONTS supernodes each (a combined factorisation and solve sweeps the levels
    // below them beside their factorisation). Measured on the 500 × 500 plate (one combined call, FP64): 41.5 ms with
    // the levels of one supernode, 42.0 ms with those of up to 4 (the sweep then slows the 4 fronts' update
    // matrices more than it gains), 42.7 ms without the overlap.
    constexpr int32_t OVERLAP_FRONTS = 1;
    p.overlapLevel = -1;
    for (int32_t l = p.levels - 1; l >= 1 && !p.outOfCore && p.levelStart[l + 1] - p.levelStart[l] <= OVERLAP_FRONTS;
         --l)
        p.overlapLevel = l;
    for (int32_t l = 0; l < p.levels; ++l)
    {
        const size_t firstLaunch = p.schedule.size();
        struct LevelMark
        {
            oosg_plan& plan;
            size_t from;
            int32_t level;
            ~LevelMark()
            {
                for (size_t q = from; q < plan.schedule.size(); ++q)
                    plan.schedule[q].level = level;
            }
        } mark{p, firstLaunch, l};
        std::vector<int32_t> large;
        // Small fronts by order bucket, one launch each: a launch's LDS is sized for its largest front, so the
        // buckets of smaller fronts fit more workgroups on a CU.
        constexpr int BUCKETS[] = {48, 64, 80, oosg::SMALL_FRONT_MAX};
        std::vector<int32_t> bucket[4];
        for (int32_t q = p.levelStart[l]; q < p.levelStart[l + 1]; ++q)
        {
            const int32_t s = p.levelList[q];
            const int order = p.ncol[s] + p.nrow[s];
            if (order <= oosg::SMALL_FRONT_MAX)
            {
                int b = 0;
                while (order > BUCKETS[b])
                    ++b;
                bucket[b].push_back(s);
            }
            else
                large.push_back(s);
        }
        for (int b = 0; b < 4; ++b)
        {
            if (bucket[b].empty())
                continue;
            Launch launch{LaunchKind::Small, PROFILE_SMALL};
            launch.offset = static_cast<int64_t>(p.smallList.size());
            launch.count = static_cast<int>(bucket[b].size());
            for (const int32_t s : bucket[b])
            {
                p.smallList.push_back(s);
                launch.maxOrder = std::max(launch.maxOrder, p.ncol[s] + p.nrow[s]);
                launch.flops += frontFlops(p.ncol[s], p.nrow[s]);
            }
            p.smallFronts += launch.count;
            p.schedule.push_back(launch);
        }
        p.largeFronts += static_cast<int64_t>(large.size());
        if (large.empty())
            continue;
        // Assembly blocks, each with the parts of the children's update matrices that land in it.
        const int64_t blockStart = static_cast<int64_t>(p.blocks.size());
        int maxC = 0;
        for (const int32_t s : large)
        {
            appendPanelBlocks(p, s, p.blocks, p.items);
            maxC = std::max(maxC, p.ncol[s]);
        }
        Launch assemble{LaunchKind::Assemble, PROFILE_ASSEMBLE};
        assemble.offset = blockStart;
        assemble.count = static_cast<int>(p.blocks.size() - blockStart);
        for (size_t x = static_cast<size_t>(blockStart); x < p.blocks.size(); ++x)
        {
            const oosg::AssemblyBlock& blk = p.blocks[x];
            for (int32_t j = blk.j0; j < blk.j1; ++j)
                assemble.values += std::max(0, blk.i1 - std::max(blk.i0, j));
            for (int32_t it = blk.itemStart; it < blk.itemEnd; ++it)
            {
                const oosg::AssemblyItem& item = p.items[it];
                for (int32_t jj = item.q0; jj < item.q1; ++jj)
                    assemble.values += std::max(0, item.ib - std::max(jj, item.ia));
            }
        }
        p.schedule.push_back(assemble);
        // Panel steps.
        // A record of EVENT_SOLVED on the critical stream delays its next launch by several microseconds, so a step
        // records only when the bulk update needs it, every SOLVED_RECORD_STEPS steps, and at the level's last step.
        // pendingInverses: the whole inverses' steps whose panel solve is done but not yet recorded (they wait for
        // the next record).
        std::vector<size_t> pendingInverses;
        const auto launchPendingInverses = [&]
        {
            for (const size_t K : pendingInverses)
            {
                if (p.inverseRounds[l][K].rowCount > 0)
                    p.schedule.push_back(inverseRowsLaunch(p.inverseRounds[l][K], EVENT_SOLVED));
                p.schedule.push_back(inverseLaunch(p.inverseRounds[l][K], EVENT_SOLVED));
            }
            pendingInverses.clear();
        };
        for (int k0 = 0; k0 < maxC; k0 += PANEL_BLOCK)
        {
            Launch diagonal{LaunchKind::Diagonal, PROFILE_DIAGONAL};
            diagonal.k0 = k0;
            diagonal.offset = static_cast<int64_t>(p.stepFronts.size());
            Launch solve{LaunchKind::Tiles, PROFILE_PANEL_SOLVE};
            solve.offset = static_cast<int64_t>(p.tiles.size());
            for (const int32_t s : large)
            {
                const int c = p.ncol[s], n = c + p.nrow[s];
                if (c <= k0)
                    continue;
                const int kb = std::min(PANEL_BLOCK, c - k0), k1 = k0 + kb;
                p.stepFronts.push_back(s);
                diagonal.flops += frontFlops(kb, 0);
                // L(rows, blk) = F(rows, blk)·L11⁻ᵀ·D⁻¹: one mode-1 tile per 64 rows below the block.
                for (int row = k1; row < n; row += 64)
                {
                    oosg::UpdateTile t{};
                    t.aOff = p.panelOff[s] + row + int64_t(k0) * n;
                    t.cOff = t.aOff;
                    t.bOff = p.linvOff[s] + int64_t(k0) * PANEL_BLOCK;
                    t.dOff = p.panelOff[s] + int64_t(k0) * (n + 1);
                    t.ld = n;
                    t.ldb = kb;
                    t.ldc = n;
                    t.m = std::min(64, n - row);
                    t.n = kb;
                    t.k = kb;
                    t.mode = 1;
                    p.tiles.push_back(t);
                    solve.flops += 2.0 * t.m * (static_cast<double>(kb) * (kb + 1) / 2.0);
                }
            }
            solve.count = static_cast<int>(p.tiles.size() - solve.offset);
            // The trailing update, blocked twice: a step updates only the columns of its own super-block of
            // SUPER_BLOCK columns (K = 64); after the super-block's last step, one product with K = SUPER_BLOCK
            // updates the columns after it: the next super-block on the critical stream (its steps need them), the
            // columns after that on the bulk stream, beside the next super-block's steps.
            const int k1Step = k0 + PANEL_BLOCK;
            const int super0 = k0 - k0 % SUPER_BLOCK, super1 = super0 + SUPER_BLOCK;
            // `finishes`: the first column of the block whose diagonal tile, in this update, gets its last
            // contribution (-1: none).
            const auto addUpdates = [&](Launch& update, int colFrom, int colTo, int t0, int t1, int finishes)
            {
                update.offset = static_cast<int64_t>(p.tiles.size());
                for (const int32_t s : large)
                {
                    const int c = p.ncol[s], n = c + p.nrow[s];
                    const int from = colFrom, to = std::min(c, colTo);
                    const int termsEnd = std::min(c, t1);
                    if (termsEnd <= t0 || from >= to)
                        continue;
                    for (int J0 = from; J0 < to; J0 += 64)
                        for (int I0 = J0; I0 < n; I0 += 64)
                        {
                            oosg::UpdateTile t{};
                            t.cOff = p.panelOff[s] + I0 + int64_t(J0) * n;
                            t.ldc = n;
                            t.aOff = p.panelOff[s] + I0 + int64_t(t0) * n;
                            t.bOff = p.panelOff[s] + J0 + int64_t(t0) * n;
                            t.dOff = p.panelOff[s] + int64_t(t0) * (n + 1);
                            t.ld = n;
                            t.m = std::min(64, n - I0);
                            t.n = std::min(64, to - J0);
                            t.k = termsEnd - t0;
                            t.target = 0;
                            t.lower = I0 == J0;
                            if (I0 == J0 && J0 == finishes)
                            {
                                // This update finishes the next step's diagonal block: factorise it here.
                                t.mode = 2;
                                t.supernode = s;
                                t.rowBase = J0;
                                t.colBase = std::min(PANEL_BLOCK, c - J0);
                                update.flops += frontFlops(t.colBase, 0);
                                update.fuses = true;
                                update.profile = PROFILE_PANEL_UPDATE_FUSED;
                            }
                            p.tiles.push_back(t);
                            update.flops += tileFlops(t);
                        }
                }
                update.count = static_cast<int>(p.tiles.size() - update.offset);
            };
            Launch inner{LaunchKind::Tiles, PROFILE_PANEL_UPDATE}, next{LaunchKind::Tiles, PROFILE_PANEL_UPDATE},
                rest{LaunchKind::Tiles, PROFILE_PANEL_UPDATE};
            addUpdates(inner, k1Step, super1, k0, k1Step, k1Step < super1 ? k1Step : -1);
            if (k1Step == super1)
            {
                addUpdates(next, super1, super1 + SUPER_BLOCK, super0, super1, super1);
                addUpdates(rest, super1 + SUPER_BLOCK, INT32_MAX, super0, super1, -1);
            }
            diagonal.count = static_cast<int>(p.stepFronts.size() - diagonal.offset);
            diagonal.lane = solve.lane = inner.lane = next.lane = LANE_CRITICAL;
            const bool records = rest.count > 0 || k0 + PANEL_BLOCK >= maxC ||
                                 (k0 / PANEL_BLOCK) % SOLVED_RECORD_STEPS == SOLVED_RECORD_STEPS - 1;
            Launch& solvedAt = solve.count > 0 ? solve : diagonal;
            if (records)
                solvedAt.record = EVENT_SOLVED;
            // From the second step on, the previous step's update factorised this step's diagonal blocks.
            if (k0 == 0)
                p.schedule.push_back(diagonal);
            if (solve.count > 0)
                p.schedule.push_back(solve);
            if (inner.count > 0)
                p.schedule.push_back(inner);
            if (next.count > 0)
            {
                next.wait = EVENT_RESTED; // the previous super-block's bulk update wrote these columns too
                p.schedule.push_back(next);
            }
            if (rest.count > 0)
            {
                rest.lane = LANE_BULK;
                rest.wait = EVENT_SOLVED;
                rest.record = EVENT_RESTED;
                p.schedule.push_back(rest);
            }
            // Step K = k0 / 64 of the level's whole inverses reads the panels' column block K (rows below block K),
            // final once this step's solve is done: it runs beside the next steps instead of after the level.
            const size_t K = static_cast<size_t>(k0 / PANEL_BLOCK);
            const std::vector<oosg_plan::InverseRound>& rounds = p.inverseRounds[l];
            if (K + 1 < rounds.size() && rounds[K].count > 0)
            {
                if (solve.count == 0)
                    throw std::logic_error("plan: an inverse step without a panel solve");
                pendingInverses.push_back(K);
            }
            // A recorded step's solve covers the earlier steps' (one stream, in order).
            if (records && (solve.count > 0 || k0 == 0))
                launchPendingInverses();
        }
        if (!pendingInverses.empty())
        {
            p.schedule.push_back(syncLaunch(LANE_CRITICAL, -1, EVENT_SOLVED));
            launchPendingInverses();
        }
        // The level's panels are final once the bulk stream is done with them.
        p.schedule.push_back(syncLaunch(LANE_CRITICAL, EVENT_RESTED, EVENT_PANELS));
        // The update matrices: U(r × r) = −L_R·D·L_Rᵀ + the children's parts, K = c. Each 64 × 64 quarter of a tile
        // lists its children's parts (the assembly items of its rows and columns).
        const auto addQuarterItems = [&p](int32_t s, int i0, int i1, int j0, int j1)
        {
            for (int32_t q = p.childStart[s]; q < p.childStart[s + 1]; ++q)
            {
                const int32_t ch = p.children[q];
                const int32_t* rel = p.rel.data() + p.rowStart[ch];
                const int32_t rc = p.nrow[ch];
                const auto at = [&](int value)
                { return static_cast<int32_t>(std::lower_bound(rel, rel + rc, value) - rel); };
                const oosg::AssemblyItem it{ch, at(j0), at(j1), at(i0), at(i1), 0};
                if (it.q0 < it.q1 && std::max(it.q0, it.ia) < it.ib)
                    p.extendItems.push_back(it);
            }
        };
        const auto addUpdateTile = [&](std::vector<oosg::UpdateTile>& list, int32_t s, int I0, int J0, int size)
        {
            const int c = p.ncol[s], r = p.nrow[s], n = c + r;
            oosg::UpdateTile t{};
            t.cOff = p.updateOff[s] + I0 + int64_t(J0) * r;
            t.ldc = r;
            t.aOff = p.panelOff[s] + c + I0;
            t.bOff = p.panelOff[s] + c + J0;
            t.dOff = p.panelOff[s];
            t.ld = n;
            t.m = std::min(size, r - I0);
            t.n = std::min(size, r - J0);
            t.k = c;
            t.target = 1;
            t.lower = I0 == J0;
            t.supernode = s;
            t.rowBase = c + I0;
            t.colBase = c + J0;
            t.items = static_cast<int32_t>(p.itemStarts.size());
            for (int qi = 0; qi < size; qi += 64)
                for (int qj = 0; qj < size; qj += 64)
                {
                    p.itemStarts.push_back(static_cast<int32_t>(p.extendItems.size()));
                    const int qm = std::min(64, t.m - qi), qn = std::min(64, t.n - qj);
                    if (qm > 0 && qn > 0 && !(t.lower && qi + qm - 1 < qj))
                        addQuarterItems(s, c + I0 + qi, c + I0 + qi + qm, c + J0 + qj, c + J0 + qj + qn);
                }
            p.itemStarts.push_back(static_cast<int32_t>(p.extendItems.size()));
            list.push_back(t);
        };
        const int64_t tileStart = static_cast<int64_t>(p.tiles.size());
        for (const int32_t s : large)
        {
            const int r = p.nrow[s];
            for (int J0 = 0; J0 < r; J0 += 64)
                for (int I0 = J0; I0 < r; I0 += 64)
                    addUpdateTile(p.tiles, s, I0, J0, 64);
        }
        if (static_cast<int64_t>(p.tiles.size()) > tileStart)
        {
            Launch tilesLaunch{LaunchKind::Tiles, PROFILE_UPDATE_MATRIX};
            tilesLaunch.offset = tileStart;
            tilesLaunch.count = static_cast<int>(p.tiles.size() - tileStart);
            for (int64_t q = tileStart; q < static_cast<int64_t>(p.tiles.size()); ++q)
                tilesLaunch.flops += tileFlops(p.tiles[q]);
            // The same products on 128 × 128 tiles.
            tilesLaunch.bigOffset = static_cast<int64_t>(p.bigTiles.size());
            for (const int32_t s : large)
            {
                const int r = p.nrow[s];
                for (int J0 = 0; J0 < r; J0 += 128)
                    for (int I0 = J0; I0 < r; I0 += 128)
                        addUpdateTile(p.bigTiles, s, I0, J0, 128);
            }
            tilesLaunch.bigCount = static_cast<int>(p.bigTiles.size() - tilesLaunch.bigOffset);
            // Before measurement: big tiles where they still fill the chip several times over.
            tilesLaunch.big[0] = tilesLaunch.big[1] = tilesLaunch.bigCount >= 2048;
            p.schedule.push_back(tilesLaunch);
        }
        // The last launch of the level's whole inverses (X_IJ = L_II⁻¹·T_IJ), once its panels are final.
        if (!p.inverseRounds[l].empty() && p.inverseRounds[l].back().count > 0)
            p.schedule.push_back(inverseLaunch(p.inverseRounds[l].back(), EVENT_PANELS));
    }
    // The other streams join the critical one (each first waits for a record of this sequence, so that both belong to
    // it even when they had nothing to do).
    p.schedule.push_back(syncLaunch(LANE_CRITICAL, -1, EVENT_FORK));
    p.schedule.push_back(syncLaunch(LANE_BULK, EVENT_FORK, EVENT_JOIN_BULK));
    p.schedule.push_back(syncLaunch(LANE_INVERSES, EVENT_FORK, EVENT_JOIN_INVERSES));
    p.schedule.push_back(syncLaunch(LANE_CRITICAL, EVENT_JOIN_BULK, -1));
    p.schedule.push_back(syncLaunch(LANE_CRITICAL, EVENT_JOIN_INVERSES, -1));
}

// The arrays the sweeps and the whole inverses read, and the device view of the structure.
void uploadSweepPlan(oosg_plan& p)
{
    const hipStream_t st = p.stream;
    p.dFirst = upload(p.first, st, "plan: first columns");
    p.dNcol = upload(p.ncol, st, "plan: column counts");
    p.dNrow = upload(p.nrow, st, "plan: row counts");
    p.dPanelOff = upload(p.panelOff, st, "plan: panel offsets");
    p.dUpdateOff = upload(p.updateOff, st, "plan: update offsets");
    p.dRowStart = upload(p.rowStart, st, "plan: row starts");
    p.dRows = upload(p.rows, st, "plan: rows");
    p.dRel = upload(p.rel, st, "plan: relative indices");
    p.dChildStart = upload(p.childStart, st, "plan: child starts");
    p.dChildren = upload(p.children, st, "plan: children");
    p.dLinvOff = upload(p.linvOff, st, "plan: inverse offsets");
    p.dForwardTiles = upload(p.forwardTiles, st, "plan: forward row tiles");
    p.dBackwardBlocks = upload(p.backwardBlocks, st, "plan: backward column blocks");
    p.dForwardSubtractions = upload(p.forwardSubtractions, st, "plan: forward subtractions");
    p.dBackwardSubtractions = upload(p.backwardSubtractions, st, "plan: backward subtractions");
    p.dForwardTriangles = upload(p.forwardTriangles, st, "plan: triangle blocks");
    p.dBackwardTriangles = upload(p.backwardTriangles, st, "plan: triangle blocks");
    p.dForwardTriangleSums = upload(p.forwardTriangleSums, st, "plan: triangle blocks");
    p.dBackwardTriangleSums = upload(p.backwardTriangleSums, st, "plan: triangle blocks");
    p.dInverseItems = upload(p.inverseItems, st, "plan: triangle inverse blocks");
    p.dLevelList = upload(p.levelList, st, "plan: levels");
    p.dTinyList = upload(p.tinyList, st, "plan: tiny supernodes");
    p.dOtherList = upload(p.otherList, st, "plan: sweep supernodes");
    p.dMediumList = upload(p.mediumList, st, "plan: sweep supernodes");
    p.dPullBase = upload(p.pullBase, st, "plan: pull table");
    p.dOtherPosStart = upload(p.otherPosStart, st, "plan: gather lists");
    p.dGatherStart = upload(p.gatherStart, st, "plan: gather lists");
    p.dGatherRows = upload(p.gatherRows, st, "plan: gather lists");
    p.dGatherChunks = upload(p.gatherChunks, st, "plan: gather lists");
    p.dPulls = upload(p.pulls, st, "plan: pull table");
    p.dp.first = p.dFirst.as<int32_t>();
    p.dp.ncol = p.dNcol.as<int32_t>();
    p.dp.nrow = p.dNrow.as<int32_t>();
    p.dp.panelOff = p.dPanelOff.as<int64_t>();
    p.dp.updateOff = p.dUpdateOff.as<int64_t>();
    p.dp.rowStart = p.dRowStart.as<int64_t>();
    p.dp.rows = p.dRows.as<int32_t>();
    p.dp.rel = p.dRel.as<int32_t>();
    p.dp.childStart = p.dChildStart.as<int32_t>();
    p.dp.children = p.dChildren.as<int32_t>();
    p.dp.linvOff = p.dLinvOff.as<int64_t>();
    p.dp.pullBase = p.dPullBase.as<int32_t>();
    p.dp.pulls = p.dPulls.as<int4>();
    p.dp.pullWidth = p.pullWidth;
}

void uploadPlan(oosg_plan& p)
{
    uploadSweepPlan(p);
    const hipStream_t st = p.stream;
    p.dSmallList = upload(p.smallList, st, "plan: small fronts");
    p.dStepFronts = upload(p.stepFronts, st, "plan: panel step fronts");
    p.dBlocks = upload(p.blocks, st, "plan: assembly blocks");
    p.dItems = upload(p.items, st, "plan: assembly items");
    p.dTiles = upload(p.tiles, st, "plan: update tiles");
    p.dBigTiles = upload(p.bigTiles, st, "plan: big update tiles");
    p.dItemStarts = upload(p.itemStarts, st, "plan: update tile items");
    p.dExtendItems = upload(p.extendItems, st, "plan: update tile items");
    p.dTrailingBlocks = upload(p.trailingBlocks, st, "plan: trailing front");
    p.dTrailingItems = upload(p.trailingItems, st, "plan: trailing front");
    p.dTrailingPullStart = upload(p.trailingPullStart, st, "plan: trailing front");
    p.dTrailingPulls = upload(p.trailingPulls, st, "plan: trailing front");
    allocateCounted(p, p.updates, static_cast<size_t>(p.updateOff[p.S]) * sizeof(double), "plan: update matrices");
    check(hipStreamSynchronize(st), "plan: upload");
}

// The plan's own arrays on the device (structure, launch lists), in bytes.
uint64_t structureBytes(const oosg_plan& p)
{
    return (p.first.size() * 3 + p.rows.size() * 2 + p.children.size() + p.smallList.size() + p.stepFronts.size() +
            p.levelList.size() * 2 + p.mediumList.size() + p.pullBase.size() + p.pulls.size() + p.gatherStart.size() +
            p.gatherRows.size() + p.otherPosStart.size() * 2 + p.gatherChunks.size() * 8) *
               4 +
           p.panelOff.size() * 24 +
           (p.forwardTiles.size() + p.backwardBlocks.size() + p.forwardSubtractions.size() +
            p.backwardSubtractions.size()) *
               16 +
           (p.forwardTriangles.size() + p.backwardTriangles.size() + 2 * p.forwardTriangleSums.size()) * 16 +
           p.inverseItems.size() * 8 + p.blocks.size() * sizeof(oosg::AssemblyBlock) +
           p.items.size() * sizeof(oosg::AssemblyItem) +
           (p.tiles.size() + p.bigTiles.size()) * sizeof(oosg::UpdateTile) + p.itemStarts.size() * 4 +
           p.extendItems.size() * sizeof(oosg::AssemblyItem) +
           p.trailingBlocks.size() * sizeof(oosg::AssemblyBlock) +
           p.trailingItems.size() * sizeof(oosg::AssemblyItem) +
           (p.trailingPullStart.size() + p.trailingPulls.size()) * 4;
}

uint64_t factorStorageBytes(const oosg_plan& p)
{
    return (static_cast<uint64_t>(std::max<int64_t>(p.panelOff[p.S], 1)) +
            static_cast<uint64_t>(std::max<int64_t>(p.linvOff[p.S], 1)) + 2 * static_cast<uint64_t>(std::max(p.n, 1))) *
           sizeof(double);
}

// The sizes of the vectors of the solves and refinements (16 right-hand sides), in the order reserveSolveBuffers
// allocates them.
std::vector<uint64_t> solveBufferBytes(const oosg_plan& p)
{
    const uint64_t vector = static_cast<uint64_t>(std::max(p.n, 1)) * 16 * sizeof(double);
    return {vector,                                                                         // solveX
            static_cast<uint64_t>(std::max<int64_t>(p.rowStart[p.S], 1)) * 16 * sizeof(double), // solveU
            static_cast<uint64_t>(std::max<int64_t>(p.partialSlots, 1)) * 64 * 16 * sizeof(double), // solvePartials
            3 * 16 * sizeof(unsigned long long),                                            // solveNorms
            vector, vector, vector, vector, vector,                                         // cgB .. cgQ
            static_cast<uint64_t>((oosg::GRAM_GROUPS + 16) * oosg::GRAM_PAIRS * 256) * sizeof(double), // cgPartial
            static_cast<uint64_t>(oosg::GRAM_PAIRS) * 256 * sizeof(double),                  // cgDots
            static_cast<uint64_t>(std::max(p.S, 1)) * sizeof(int32_t),                      // pivotStatus
            sizeof(oosg::PivotStatistics),                                                  // pivotStats
            static_cast<uint64_t>(std::max(p.S, 1)) * sizeof(float)};                       // supernodeGrowth
}

// Everything a plan reserves: its arrays, the update matrices, one factor's storage, the solves' vectors.
uint64_t planBytes(const oosg_plan& p)
{
    uint64_t bytes = structureBytes(p) + static_cast<uint64_t>(std::max<int64_t>(p.updateOff[p.S], 1)) * 8 +
                     factorStorageBytes(p);
    for (const uint64_t b : solveBufferBytes(p))
        bytes += b;
    return bytes;
}

void reserveSolveBuffers(oosg_plan& p)
{
    const std::vector<uint64_t> bytes = solveBufferBytes(p);
    DeviceBuffer* buffers[] = {&p.solveX, &p.solveU, &p.solvePartials, &p.solveNorms, &p.cgB, &p.cgX,
                               &p.cgR,    &p.cgP,    &p.cgQ,          &p.cgPartial,   &p.cgDots,
                               &p.pivotStatus, &p.pivotStats, &p.supernodeGrowth};
    for (size_t k = 0; k < bytes.size(); ++k)
        ensureCounted(p, *buffers[k], bytes[k], "plan: solve vectors");
}

// ---- factorisation -------------------------------------------------------------------------------------------------

template <typename TP>
void forwardSweep(oosg_plan& p, const TP* panels, const TP* linv, int k, int32_t l0, int32_t l1, hipStream_t st);
void ensureSolveBuffers(oosg_plan& p);

// cancellable: launched one launch at a time (not captured), so the cancellation is polled between levels.
template <typename T>
void runFactor(oosg_plan& p, const oosg_matrix& m, T* panels, T* linv, double* diagOut, int32_t* zeroAt, bool scaled,
               bool cancellable = false, void* spool = nullptr, int32_t levelFrom = 0,
               int32_t levelTo = std::numeric_limits<int32_t>::max())
{
    // A part of the levels (a mixed factor's) runs on the plan's stream alone, launch by launch.
    const bool part = levelFrom > 0 || levelTo < p.levels;
    if (p.outOfCore && spool == nullptr)
        fail(OOSG_INTERNAL, "factor: an out-of-core factorisation without its spool");
    const hipStream_t st = p.stream;
    oosg::DeviceMatrix a{m.dOffsets.as<int64_t>(), m.dRows.as<int32_t>(), m.dValues.as<double>(),
                         m.dPosition.as<int32_t>(), m.dColumn.as<int32_t>(),
                         scaled ? m.dScale.as<double>() : nullptr};
    T* updates = p.updates.as<T>();
    const int32_t* smallList = p.dSmallList.as<int32_t>();
    const int32_t* stepFronts = p.dStepFronts.as<int32_t>();
    const oosg::AssemblyBlock* blocks = p.dBlocks.as<oosg::AssemblyBlock>();
    const oosg::AssemblyItem* items = p.dItems.as<oosg::AssemblyItem>();
    const int2* inverseItems = p.dInverseItems.as<int2>();
    const oosg::UpdateTile* tiles = p.dTiles.as<oosg::UpdateTile>();
    const bool timing = p.timeLaunches;
    // The sweep ahead (a combined factorisation and solve): its levels are launched one at a time between the
    // factorisation's launches, AHEAD_SPREAD of them apart, so that the host's launches of the factorisation's top
    // levels do not wait behind all of the sweep's (the GPU runs both: the sweep's stream waits for the levels below
    // aheadLevel only).
    constexpr int AHEAD_SPREAD = 4;
    bool aheadIssued = false;
    int32_t aheadNext = 0; // the next level of the sweep to launch
    int aheadWait = 0;     // factorisation launches before it
    const auto aheadLevelLaunch = [&]()
    {
        forwardSweep<T>(p, panels, linv, p.aheadK, aheadNext, aheadNext + 1, p.sweepStream);
        if (++aheadNext == p.aheadLevel)
            check(hipEventRecord(p.aheadEvents[LANES], p.sweepStream), "factor: sweep ahead");
        aheadWait = AHEAD_SPREAD;
    };
    if (timing)
        while (p.launchEvents.size() < 2 * p.schedule.size())
        {
            hipEvent_t e = nullptr;
            check(hipEventCreate(&e), "hipEventCreate");
            p.launchEvents.push_back(e);
        }
    // Out of core: the launches of level l write window l mod 2 of the panels and inverses; once a level's launches
    // are queued, its windows are copied to the spool on the copy stream, and a window is written again only after
    // that copy (oocFree).
    T* const panelBase = panels;
    T* const linvBase = linv;
    int32_t current = -1;
    const auto spoolLevel = [&](int32_t level)
    {
        const int w = level % 2;
        check(hipEventRecord(p.oocReady[w], st), "factor: spool");
        check(hipStreamWaitEvent(p.copyStream, p.oocReady[w], 0), "factor: spool");
        const int64_t values = p.spoolPanelBase[level + 1] - p.spoolPanelBase[level];
        const int64_t inverses = p.spoolLinvBase[level + 1] - p.spoolLinvBase[level];
        if (values > 0)
            check(hipMemcpyAsync(static_cast<T*>(spool) + p.spoolPanelBase[level], panelBase + w * p.windowPanels,
                                 static_cast<size_t>(values) * sizeof(T), hipMemcpyDeviceToHost, p.copyStream),
                  "factor: spool");
        if (inverses > 0)
            check(hipMemcpyAsync(static_cast<T*>(spool) + p.spoolPanelBase[p.levels] + p.spoolLinvBase[level],
                                 linvBase + w * p.windowLinv, static_cast<size_t>(inverses) * sizeof(T),
                                 hipMemcpyDeviceToHost, p.copyStream),
                  "factor: spool");
        check(hipEventRecord(p.oocFree[w], p.copyStream), "factor: spool");
    };
    for (size_t li = 0; li < p.schedule.size(); ++li)
    {
        const Launch& l = p.schedule[li];
        if (part && (l.kind == LaunchKind::Sync || l.level < levelFrom || l.level >= levelTo))
            continue;
        if (cancellable && li > 0 && l.level != p.schedule[li - 1].level)
            pollCancellation(p);
        if (p.outOfCore && l.kind != LaunchKind::Sync && l.level != current)
        {
            if (l.level < current)
                fail(OOSG_INTERNAL, "factor: the launches of an out-of-core plan go back a level");
            if (current >= 0)
                spoolLevel(current);
            current = l.level;
            // The window's previous level copied out first (no wait before its first record).
            check(hipStreamWaitEvent(st, p.oocFree[current % 2], 0), "factor: spool");
            panels = panelBase + (current % 2) * p.windowPanels;
            linv = linvBase + (current % 2) * p.windowLinv;
        }
        const bool concurrent = p.concurrent[sizeof(T) == sizeof(float) ? OOSG_FP32 : OOSG_FP64] &&
                                !p.profiler.enabled && !p.outOfCore && !part;
        if (p.aheadLevel > 0 && !aheadIssued && l.level >= p.aheadLevel)
        {
            // Everything launched so far is the factorisation of the levels below aheadLevel, on whichever streams
            // it ran: the sweep stream waits for all of it, then sweeps those levels forward.
            aheadIssued = true;
            const int streams = concurrent ? LANES : 1;
            for (int q = 0; q < streams; ++q)
            {
                check(hipEventRecord(p.aheadEvents[q], q == 0 ? st : p.lanes[q]), "factor: sweep ahead");
                check(hipStreamWaitEvent(p.sweepStream, p.aheadEvents[q], 0), "factor: sweep ahead");
            }
            aheadLevelLaunch();
        }
        else if (aheadIssued && aheadNext < p.aheadLevel && l.kind != LaunchKind::Sync && --aheadWait <= 0)
            aheadLevelLaunch();
        const bool lanes = concurrent && p.lanes[l.lane] != nullptr;
        const hipStream_t ls = lanes ? p.lanes[l.lane] : st;
        if (lanes && l.wait >= 0)
            check(hipStreamWaitEvent(ls, p.laneEvents[l.wait], 0), "factor: stream order");
        if (l.kind == LaunchKind::Sync)
        {
            if (lanes && l.record >= 0)
                check(hipEventRecord(p.laneEvents[l.record], ls), "factor: stream order");
            if (timing)
            {
                check(hipEventRecord(p.launchEvents[2 * li], ls), "hipEventRecord");
                check(hipEventRecord(p.launchEvents[2 * li + 1], ls), "hipEventRecord");
            }
            continue;
        }
        if (timing)
            check(hipEventRecord(p.launchEvents[2 * li], ls), "hipEventRecord");
        const int precision = sizeof(T) == sizeof(float) ? OOSG_FP32 : OOSG_FP64;
        const bool big = l.kind == LaunchKind::Tiles && l.bigCount > 0 &&
                         (p.tilePolicy == 1 || (p.tilePolicy == 2 && l.big[precision]));
        const int slot = p.profiler.begin(big ? PROFILE_UPDATE_MATRIX_BIG : l.profile,
                                          l.kind == LaunchKind::Assemble ? l.values * sizeof(T) : l.flops, ls);
        switch (l.kind)
        {
        case LaunchKind::Sync:
            break;
        case LaunchKind::Small:
            check(oosg::launchSmallFronts<T>(p.dp, a, panels, updates, linv, smallList + l.offset, l.count,
                                             l.maxOrder, diagOut, zeroAt, ls),
                  "factor: small fronts");
            break;
        case LaunchKind::Assemble:
            check(oosg::launchAssembly<T>(p.dp, a, panels, updates, blocks + l.offset, items, l.count, ls),
                  "factor: assembly");
            break;
        case LaunchKind::Diagonal:
            check(oosg::launchDiagonalBlocks<T>(p.dp, panels, linv, stepFronts + l.offset, l.count, l.k0, diagOut,
                                                zeroAt, ls),
                  "factor: diagonal blocks");
            break;
        case LaunchKind::Tiles:
        {
            if (l.profile == PROFILE_UPDATE_MATRIX)
                check(oosg::launchExtendTiles<T>(
                          p.dp, panels, updates,
                          big ? p.dBigTiles.as<oosg::UpdateTile>() + l.bigOffset : tiles + l.offset,
                          big ? l.bigCount : l.count, p.dItemStarts.as<int32_t>(),
                          p.dExtendItems.as<oosg::AssemblyItem>(), big ? 128 : 64, ls),
                      "factor: update matrices");
            else
                check(oosg::launchUpdateTiles<T>(p.dp, panels, updates, linv, tiles + l.offset, l.count, diagOut,
                                                 zeroAt, l.fuses, ls),
                      "factor: update tiles");
            break;
        }
        case LaunchKind::Inverse:
            check(oosg::launchTriangleInverse<T>(p.dp, panels, linv, inverseItems + l.offset, l.count, l.k0, ls),
                  "factor: triangle inverses");
            break;
        }
        p.profiler.end(slot, ls);
        if (lanes && l.record >= 0)
            check(hipEventRecord(p.laneEvents[l.record], ls), "factor: stream order");
        if (timing)
            check(hipEventRecord(p.launchEvents[2 * li + 1], ls), "hipEventRecord");
    }
    while (aheadIssued && aheadNext < p.aheadLevel)
        aheadLevelLaunch();
    if (p.outOfCore && current >= 0)
    {
        spoolLevel(current);
        // The caller's later work on the plan's stream follows the last copy.
        check(hipStreamWaitEvent(st, p.oocFree[current % 2], 0), "factor: spool");
    }
    if (timing)
    {
        check(hipStreamSynchronize(st), "factor: timing");
        p.launchMs.assign(p.schedule.size(), 0.0f);
        for (size_t li = 0; li < p.schedule.size(); ++li)
            check(hipEventElapsedTime(&p.launchMs[li], p.launchEvents[2 * li], p.launchEvents[2 * li + 1]),
                  "hipEventElapsedTime");
    }
}

std::string equationList(const std::vector<int32_t>& equations, size_t total)
{
    std::string text;
    const size_t listed = std::min<size_t>(equations.size(), 20);
    for (size_t k = 0; k < listed; ++k)
        text += "  equation " + std::to_string(equations[k]) + "\n";
    if (total > listed)
        text += "  and " + std::to_string(total - listed) + " more\n";
    return text;
}

// The zero-pivot test, the inertia and the ratio test of MultifrontalDriver::finish
// (src/linalg/supernodal/MultifrontalDriver.cpp:704-820), on the host from the pivots the kernels wrote.
void factoriseInto(oosg_plan* plan, const oosg_matrix* matrix, const oosg_factor_options* options, oosg_factor** out);

void holdPlan(oosg_plan* p)
{
    ++p->references;
}

void releasePlan(oosg_plan* p)
{
    if (--p->references > 0)
        return;
    (void)hipSetDevice(p->device);
    if (p->stream)
        (void)hipStreamSynchronize(p->stream);
    delete p;
    // A release reports nothing: an error of its calls must not surface as the next launch's error.
    (void)hipGetLastError();
}

// The pivots of a positive definite factor on the host, copied on first use (FP32: d_j / s_j², those of the matrix).
const std::vector<double>& hostDiagonal(const oosg_factor& f)
{
    if (!f.diag.empty() || f.plan->n == 0)
        return f.diag;
    const oosg_plan& p = *f.plan;
    const size_t n = static_cast<size_t>(p.n);
    std::vector<double> d(n);
    check(hipMemcpy(d.data(), f.dDiag.ptr, n * sizeof(double), hipMemcpyDeviceToHost), "factor: pivots");
    if (scaledPrecision(f.precision))
    {
        std::vector<double> scale(n);
        check(hipMemcpy(scale.data(), f.scale.ptr, n * sizeof(double), hipMemcpyDeviceToHost), "factor: scales");
        for (size_t j = 0; j < n; ++j)
            d[j] = d[j] / (scale[j] * scale[j]);
    }
    f.diag = std::move(d);
    return f.diag;
}

// The inverse of oosg::orderedKey.
double keyValue(unsigned long long key)
{
    const unsigned long long bits = (key >> 63) ? key & ~0x8000000000000000ull : ~key;
    return __builtin_bit_cast(double, bits);
}

// Zero pivots, then the ratio test (MultifrontalDriver::finish): the counts by sign, the largest ratio, the pivots
// that fail, from the statistics the device computed over the pivots.
void finish(const oosg_plan& p, const oosg_matrix& m, oosg_factor& f, const std::vector<int32_t>& zeroAt,
            const oosg::PivotStatistics& stats, const oosg_factor_options& options)
{
    // A supernode whose descendant failed is skipped in the reference: only the failures with no failed descendant
    // are reported (their ancestors' values are meaningless).
    std::vector<char> failedBelow(p.S, 0);
    std::vector<int32_t> zeroPivots;
    size_t zeroCount = 0;
    for (int32_t s = 0; s < p.S; ++s)
    {
        const bool failedHere = zeroAt[s] != INT32_MAX;
        if (failedHere && !failedBelow[s])
        {
            ++zeroCount;
            if (zeroPivots.size() < 1000)
                zeroPivots.push_back(p.first[s] + zeroAt[s]);
        }
        if ((failedHere || failedBelow[s]) && p.parent[s] >= 0)
            failedBelow[p.parent[s]] = 1;
    }
    // An FP32 factor's failures do not show the matrix singular: single precision may cause them.
    const bool fp32 = scaledPrecision(f.precision);
    const std::string fp32Note = "\nthe factor was computed in FP32: the matrix is not shown singular; factorise it in "
                                 "FP64";
    if (zeroCount > 0)
        failNaming(fp32 ? OOSG_PRECISION_INSUFFICIENT : OOSG_SINGULAR,
                   "the stiffness matrix is singular: " + std::to_string(zeroCount) +
                       " zero (or not finite) pivots met in the factorisation, a mechanism or an "
                       "unconstrained DOF at:\n" +
                       equationList(zeroPivots, zeroCount) +
                       "constrain these DOFs (single-point constraints) or connect them to the structure; "
                       "with a matrix that is meant to be indefinite, factorise it in indefinite mode" +
                       (fp32 ? fp32Note : std::string()),
                   zeroPivots);
    oosg_factor_info& info = f.info;
    const double maxRatio = options.max_ratio > 0 ? options.max_ratio : 1.0e7;
    info.positive = static_cast<int64_t>(stats.positive);
    info.negative = static_cast<int64_t>(stats.negative);
    info.zero = static_cast<int64_t>(stats.zero);
    info.issue_count = static_cast<int64_t>(stats.issues);
    info.first_issue_equation = stats.issues > 0 ? stats.firstIssue : -1;
    info.first_issue_negative = stats.issues > 0 ? stats.firstIssueNegative : 0;
    if (stats.largestAt != INT32_MAX)
    {
        info.largest_ratio = keyValue(stats.largestKey);
        info.largest_ratio_equation = stats.largestAt;
    }
    else
    {
        // No ratio is a number (or no equation).
        info.largest_ratio = p.eliminated > 0 ? NAN : 0.0;
        info.largest_ratio_equation = p.eliminated > 0 ? 0 : -1;
    }
    if (info.issue_count == 0)
        return;
    // The failing pivots, in elimination order (the equations in order: a supernode's columns are consecutive), from
    // the pivots on the host: the first OOSG_RECORDED_ISSUE_COUNT with their values (PivotIssue), as
    // MultifrontalDriver::finish records them.
    const std::vector<double>& d = hostDiagonal(f);
    const double* reference = options.reference_diagonal;
    std::vector<int32_t> equations;
    for (int32_t j = 0; j < p.eliminated && f.issues.size() < OOSG_RECORDED_ISSUE_COUNT; ++j)
    {
        const double diagonal = reference != nullptr && reference[j] >= 0.0 ? reference[j] : m.diagonal[j];
        const double ratio = diagonal / d[j];
        const bool negative = ratio < 0.0 || d[j] < 0.0;
        if (ratio > maxRatio || negative)
        {
            f.issues.push_back({j, negative ? 1 : 0, diagonal, d[j], ratio});
            equations.push_back(j);
        }
    }
    if (options.bailout >= 0)
        failNaming(fp32 ? OOSG_PRECISION_INSUFFICIENT : OOSG_SINGULAR,
                   "the stiffness matrix is nearly singular or not positive definite: " +
                       std::to_string(info.issue_count) +
                       " pivots fail the factor diagonal ratio test (ratio of the matrix diagonal to the "
                       "factor diagonal above " +
                       std::to_string(maxRatio) + ", or negative):\n" +
                       equationList(equations, static_cast<size_t>(info.issue_count)) +
                       (fp32 ? fp32Note : std::string()),
                   equations);
}

// ---- solves --------------------------------------------------------------------------------------------------------

// The inertia and pivot counts of a pivoted factor (MultifrontalDriver::finish, MultifrontalDriver.cpp:764-795): 1 × 1
// pivots by their sign, 2 × 2 blocks by the sign of the determinant, then of the trace. No ratio test: it belongs to
// the positive definite mode.
void finishPivoted(oosg_factor& f)
{
    const PivotedFactor& pf = *f.pivoted;
    oosg_factor_info& info = f.info;
    info.first_issue_equation = -1;
    info.first_issue_negative = 0;
    info.largest_ratio = 0.0;
    info.largest_ratio_equation = -1;
    const size_t n = pf.diag.size();
    for (size_t k = 0; k < n;)
    {
        const double pivot = pf.diag[k];
        if (k + 1 < n && pf.sub_[k] != 0.0)
        {
            const double second = pf.diag[k + 1];
            const double determinant = pivot * second - pf.sub_[k] * pf.sub_[k];
            if (determinant < 0.0)
            {
                ++info.positive;
                ++info.negative;
            }
            else if (determinant > 0.0)
                (pivot > 0.0 ? info.positive : info.negative) += 2;
            else
            {
                ++info.zero;
                const double trace = pivot + second;
                ++(trace > 0.0 ? info.positive : trace < 0.0 ? info.negative : info.zero);
            }
            k += 2;
            continue;
        }
        ++(pivot > 0.0 ? info.positive : pivot < 0.0 ? info.negative : info.zero);
        ++k;
    }
    info.delayed_pivots = pf.delayed;
    info.two_by_two_pivots = pf.twoByTwo;
}

void ensure(DeviceBuffer& b, size_t bytes, const char* what)
{
    if (b.bytes < bytes)
        b.allocate(bytes, what);
}

// ---- indefinite mode: Bunch–Kaufman fronts -------------------------------------------------------------------------

template <typename T> void uploadInto(DeviceBuffer& b, const std::vector<T>& host, hipStream_t st, const char* what)
{
    ensure(b, std::max<size_t>(host.size(), 1) * sizeof(T), what);
    if (!host.empty())
        check(hipMemcpyAsync(b.ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, st), what);
}

// Tiles of C −= A·Bᵀ over the lower part of rows [c0, N) × columns [c0, c1) of a pivoted front (element offsets fOff
// in the level's fronts and wOff in W), terms t in [t0, t1): A(i, t) = F(i, t), B(j, t) = W(j, t).
void addPivotTiles(std::vector<oosg::UpdateTile>& tiles, int64_t fOff, int64_t wOff, int N, int c0, int c1, int t0,
                   int t1)
{
    for (int J0 = c0; J0 < c1; J0 += 64)
        for (int I0 = J0; I0 < N; I0 += 64)
        {
            oosg::UpdateTile t{};
            t.aOff = fOff + I0 + int64_t(t0) * N;
            t.bOff = wOff + J0 + int64_t(t0) * N;
            t.cOff = fOff + I0 + int64_t(J0) * N;
            t.ld = N;
            t.ldb = N;
            t.ldc = N;
            t.m = std::min(64, N - I0);
            t.n = std::min(64, c1 - J0);
            t.k = t1 - t0;
            t.lower = I0 == J0;
            tiles.push_back(t);
        }
}

// The indefinite factorisation (MultifrontalDriver::eliminatePivoted, MultifrontalDriver.cpp:346-436), level by level:
// each front [own columns, the columns its children delayed (in child order), rows] is assembled, eliminated with
// Bunch–Kaufman pivoting among its candidates in blocks of PANEL_BLOCK columns, the candidates it cannot eliminate
// delayed to its parent; a root front cannot delay. The host reads each front's state after every block step (the
// trailing tiles depend on how far it got) and its permutation after the level. Then the panels (the first e columns
// of each front) are packed, the structure in pivot order is built for the sweeps, and the inverses of the unit
// lower triangles are formed.
void factorisePivoted(oosg_plan& p, const oosg_matrix& m, oosg_factor& f)
{
    const hipStream_t st = p.stream;
    const int32_t S = p.S;
    auto pf = std::make_unique<PivotedFactor>();
    pf->order.assign(S, 0);
    pf->eliminated.assign(S, 0);
    std::vector<int32_t> cand(S, 0), delayedIn(S, 0);
    std::vector<std::vector<int32_t>> frontRows(S);
    std::vector<double*> frontAt(S, nullptr);
    std::vector<std::vector<double>> dOf(S), sOf(S);
    const oosg::DeviceMatrix a{m.dOffsets.as<int64_t>(), m.dRows.as<int32_t>(), m.dValues.as<double>(),
                               m.dPosition.as<int32_t>(), m.dColumn.as<int32_t>(), nullptr};
    if (static_cast<int32_t>(p.pivotFronts.size()) < p.levels)
        p.pivotFronts.resize(p.levels);
    for (int32_t l = 0; l < p.levels; ++l)
    {
        const int32_t* list = p.levelList.data() + p.levelStart[l];
        const int count = p.levelStart[l + 1] - p.levelStart[l];
        std::vector<int64_t> fOff(count), wOff(count), cOff(count);
        int64_t fTotal = 0, wTotal = 0, cTotal = 0;
        int maxChildren = 0;
        for (int q = 0; q < count; ++q)
        {
            const int32_t s = list[q];
            int32_t d = 0;
            for (int32_t x = p.childStart[s]; x < p.childStart[s + 1]; ++x)
                d += cand[p.children[x]] - pf->eliminated[p.children[x]];
            maxChildren = std::max(maxChildren, p.childStart[s + 1] - p.childStart[s]);
            delayedIn[s] = d;
            cand[s] = p.ncol[s] + d;
            const int32_t N = cand[s] + p.nrow[s];
            pf->order[s] = N;
            fOff[q] = fTotal;
            fTotal += int64_t(N) * N;
            wOff[q] = wTotal;
            wTotal += int64_t(N) * cand[s];
            cOff[q] = cTotal;
            cTotal += cand[s];
        }
        ensureCounted(p, p.pivotFronts[l], static_cast<size_t>(std::max<int64_t>(fTotal, 1)) * sizeof(double),
               "factor: pivoted fronts");
        ensureCounted(p, p.pivotW, static_cast<size_t>(std::max<int64_t>(wTotal, 1)) * sizeof(double), "factor: W");
        ensureCounted(p, p.pivotPerm, static_cast<size_t>(std::max<int64_t>(cTotal, 1)) * sizeof(int32_t), "factor: pivots");
        ensureCounted(p, p.pivotDiag, static_cast<size_t>(std::max<int64_t>(cTotal, 1)) * sizeof(double), "factor: pivots");
        ensureCounted(p, p.pivotSub, static_cast<size_t>(std::max<int64_t>(cTotal, 1)) * sizeof(double), "factor: pivots");
        double* F0 = p.pivotFronts[l].as<double>();
        double* W0 = p.pivotW.as<double>();
        std::vector<oosg::PivotFront> fronts(count);
        std::vector<oosg::PivotState> states(count);
        std::vector<int32_t> perm0(static_cast<size_t>(cTotal));
        for (int q = 0; q < count; ++q)
        {
            const int32_t s = list[q];
            oosg::PivotFront& fr = fronts[q];
            fr.F = F0 + fOff[q];
            fr.W = W0 + wOff[q];
            fr.perm = p.pivotPerm.as<int32_t>() + cOff[q];
            fr.diag = p.pivotDiag.as<double>() + cOff[q];
            fr.sub = p.pivotSub.as<double>() + cOff[q];
            fr.N = pf->order[s];
            fr.C = cand[s];
            fr.c = p.ncol[s];
            fr.d = delayedIn[s];
            fr.first = p.first[s];
            frontAt[s] = fr.F;
            states[q] = oosg::PivotState{0, cand[s], 0, 0, 0, 0, 0, 0};
            for (int32_t i = 0; i < cand[s]; ++i)
                perm0[cOff[q] + i] = i;
        }
        uploadInto(p.pivotDesc, fronts, st, "factor: fronts");
        uploadInto(p.pivotStates, states, st, "factor: front states");
        check(hipMemcpyAsync(p.pivotPerm.ptr, perm0.data(), perm0.size() * sizeof(int32_t), hipMemcpyHostToDevice, st),
              "factor: pivots");
        const oosg::PivotFront* desc = p.pivotDesc.as<oosg::PivotFront>();
        oosg::PivotState* dStates = p.pivotStates.as<oosg::PivotState>();
        check(oosg::launchFillDouble(F0, fTotal, fTotal, 1, 0.0, st), "factor: fronts");
        check(oosg::launchPivotMatrix(a, desc, count, st), "factor: matrix entries");
        // The children's update matrices, the first child of every front, then the second, and so on.
        std::vector<int32_t> maps;
        std::vector<oosg::ExtendItem> items;
        std::vector<int64_t> itemMap;
        std::vector<int> rankStart(static_cast<size_t>(maxChildren) + 1, 0);
        for (int rank = 0; rank < maxChildren; ++rank)
        {
            rankStart[rank] = static_cast<int>(items.size());
            for (int q = 0; q < count; ++q)
            {
                const int32_t s = list[q];
                if (p.childStart[s] + rank >= p.childStart[s + 1])
                    continue;
                int32_t delayedStart = p.ncol[s];
                for (int32_t x = p.childStart[s]; x < p.childStart[s] + rank; ++x)
                    delayedStart += cand[p.children[x]] - pf->eliminated[p.children[x]];
                const int32_t ch = p.children[p.childStart[s] + rank];
                const int32_t Nc = pf->order[ch], ec = pf->eliminated[ch], M = Nc - ec;
                if (M == 0)
                    continue;
                const int64_t mapOff = static_cast<int64_t>(maps.size());
                for (int32_t i = 0; i < cand[ch] - ec; ++i)
                    maps.push_back(delayedStart + i);
                for (int32_t i = 0; i < p.nrow[ch]; ++i)
                {
                    const int32_t rel = p.rel[p.rowStart[ch] + i];
                    maps.push_back(rel < p.ncol[s] ? rel : rel + delayedIn[s]);
                }
                const double* src = frontAt[ch] + ec + int64_t(ec) * Nc;
                for (int32_t j0 = 0; j0 < M;)
                {
                    const int32_t j1 = std::min(M, j0 + std::max(1, 16384 / (M - j0)));
                    items.push_back(oosg::ExtendItem{src, frontAt[s], nullptr, Nc, pf->order[s], M, j0, j1, 0});
                    itemMap.push_back(mapOff);
                    j0 = j1;
                }
            }
        }
        rankStart[maxChildren] = static_cast<int>(items.size());
        if (!items.empty())
        {
            uploadInto(p.pivotMaps, maps, st, "factor: extend-add positions");
            for (size_t x = 0; x < items.size(); ++x)
                items[x].map = p.pivotMaps.as<int32_t>() + itemMap[x];
            uploadInto(p.pivotItems, items, st, "factor: extend-add items");
            for (int rank = 0; rank < maxChildren; ++rank)
                check(oosg::launchPivotExtendAdd(p.pivotItems.as<oosg::ExtendItem>() + rankStart[rank],
                                                 rankStart[rank + 1] - rankStart[rank], st),
                      "factor: extend-add");
        }
        // Block steps until every front of the level is done, each followed by its trailing tiles.
        std::vector<int32_t> active(count), next;
        for (int q = 0; q < count; ++q)
            active[q] = q;
        std::vector<int32_t> kNow(count, 0);
        std::vector<oosg::UpdateTile> tiles;
        // Every step of a front eliminates PANEL_BLOCK columns or ends a pass, and a pass that eliminates nothing
        // ends the front: this bounds the steps.
        int32_t maxCand = 0;
        for (int q = 0; q < count; ++q)
            maxCand = std::max(maxCand, cand[list[q]]);
        int64_t stepsLeft = 2 * int64_t(maxCand) + 8;
        while (!active.empty())
        {
            if (--stepsLeft < 0)
                fail(OOSG_INTERNAL, "factor: the Bunch-Kaufman steps of a level did not end");
            uploadInto(p.pivotList, active, st, "factor: active fronts");
            check(oosg::launchBunchKaufman(desc, dStates, p.pivotList.as<int32_t>(), static_cast<int>(active.size()),
                                           oosg::PANEL_BLOCK, st),
                  "factor: Bunch-Kaufman");
            check(hipMemcpyAsync(states.data(), dStates, states.size() * sizeof(oosg::PivotState),
                                 hipMemcpyDeviceToHost, st),
                  "factor: front states");
            check(hipStreamSynchronize(st), "factor: Bunch-Kaufman");
            tiles.clear();
            next.clear();
            for (const int32_t q : active)
            {
                const int32_t s = list[q];
                const int kb = kNow[q], ka = states[q].k;
                if (ka < kb || ka > cand[s])
                    fail(OOSG_INTERNAL, "factor: a front's elimination went back");
                if (ka > kb)
                    addPivotTiles(tiles, fOff[q], wOff[q], pf->order[s], ka, cand[s], kb, ka);
                kNow[q] = ka;
                if (!states[q].done)
                    next.push_back(q);
            }
            if (!tiles.empty())
            {
                uploadInto(p.pivotTiles, tiles, st, "factor: trailing tiles");
                check(oosg::launchPlainUpdateTiles(F0, W0, p.pivotTiles.as<oosg::UpdateTile>(),
                                                   static_cast<int>(tiles.size()), st),
                      "factor: trailing tiles");
            }
            active.swap(next);
        }
        // The rows below the candidates: their update matrix part −= L·Wᵀ over the eliminated columns.
        tiles.clear();
        for (int q = 0; q < count; ++q)
        {
            const int32_t s = list[q];
            const int e = states[q].k, N = pf->order[s];
            if (e > 0 && N > cand[s])
                addPivotTiles(tiles, fOff[q], wOff[q], N, cand[s], N, 0, e);
        }
        if (!tiles.empty())
        {
            uploadInto(p.pivotTiles, tiles, st, "factor: update tiles");
            check(oosg::launchPlainUpdateTiles(F0, W0, p.pivotTiles.as<oosg::UpdateTile>(),
                                               static_cast<int>(tiles.size()), st),
                  "factor: update tiles");
        }
        std::vector<int32_t> permH(static_cast<size_t>(cTotal));
        std::vector<double> dH(static_cast<size_t>(cTotal)), sH(static_cast<size_t>(cTotal));
        check(hipMemcpyAsync(permH.data(), p.pivotPerm.ptr, permH.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st),
              "factor: pivots");
        check(hipMemcpyAsync(dH.data(), p.pivotDiag.ptr, dH.size() * sizeof(double), hipMemcpyDeviceToHost, st),
              "factor: pivots");
        check(hipMemcpyAsync(sH.data(), p.pivotSub.ptr, sH.size() * sizeof(double), hipMemcpyDeviceToHost, st),
              "factor: pivots");
        check(hipStreamSynchronize(st), "factor: pivots");
        for (int q = 0; q < count; ++q)
        {
            const int32_t s = list[q];
            const int32_t e = states[q].k, C = cand[s], N = pf->order[s];
            if (p.parent[s] < 0 && e < C)
                fail(OOSG_INTERNAL, "supernodal factorisation: the root front " + std::to_string(s) + " delayed " +
                                        std::to_string(C - e) + " pivots, which a root front cannot do");
            // The equations of the front's positions before the interchanges, then after them.
            std::vector<int32_t> before;
            before.reserve(static_cast<size_t>(N));
            for (int32_t i = 0; i < p.ncol[s]; ++i)
                before.push_back(p.first[s] + i);
            for (int32_t x = p.childStart[s]; x < p.childStart[s + 1]; ++x)
            {
                const int32_t ch = p.children[x];
                before.insert(before.end(), frontRows[ch].begin() + pf->eliminated[ch],
                              frontRows[ch].begin() + cand[ch]);
            }
            before.insert(before.end(), p.rows.begin() + p.rowStart[s], p.rows.begin() + p.rowStart[s + 1]);
            std::vector<int32_t>& rows = frontRows[s];
            rows.resize(static_cast<size_t>(N));
            for (int32_t k = 0; k < C; ++k)
            {
                const int32_t from = permH[cOff[q] + k];
                if (from < 0 || from >= C)
                    fail(OOSG_INTERNAL, "factor: a candidate permutation is out of range");
                rows[k] = before[from];
            }
            for (int32_t k = C; k < N; ++k)
                rows[k] = before[k];
            dOf[s].assign(dH.begin() + cOff[q], dH.begin() + cOff[q] + e);
            sOf[s].assign(sH.begin() + cOff[q], sH.begin() + cOff[q] + e);
            pf->eliminated[s] = e;
            pf->twoByTwo += states[q].twoByTwo;
            pf->delayed += C - e;
        }
    }

    // The panels: the first e columns of each front, one contiguous range each.
    const int32_t n = p.n;
    std::vector<int64_t> panelStart(S + 1, 0);
    for (int32_t s = 0; s < S; ++s)
        panelStart[s + 1] = panelStart[s] + int64_t(pf->order[s]) * pf->eliminated[s];
    allocateCounted(p, f.panels, static_cast<size_t>(std::max<int64_t>(panelStart[S], 1)) * sizeof(double),
                    "factor: panels");
    std::vector<oosg::CopyRange> ranges;
    for (int32_t s = 0; s < S; ++s)
        for (int64_t q0 = 0; q0 < panelStart[s + 1] - panelStart[s]; q0 += 65536)
            ranges.push_back(oosg::CopyRange{frontAt[s] + q0, f.panels.as<double>() + panelStart[s] + q0,
                                             std::min<int64_t>(65536, panelStart[s + 1] - panelStart[s] - q0)});
    uploadInto(p.pivotRanges, ranges, st, "factor: panel copies");
    check(oosg::launchCopyRanges(p.pivotRanges.as<oosg::CopyRange>(), static_cast<int>(ranges.size()), st),
          "factor: panels");

    // Pivot order: the equations eliminated by supernode s get first'(s) .. first'(s) + e − 1.
    auto d = std::make_unique<oosg_plan>();
    d->device = p.device;
    d->stream = st;
    d->ownsStream = false;
    d->graphs = false;
    d->n = n;
    d->S = S;
    d->first.assign(S, 0);
    d->ncol = pf->eliminated;
    d->nrow.assign(S, 0);
    d->parent = p.parent;
    d->childStart = p.childStart;
    d->children = p.children;
    pf->pivotOf.assign(static_cast<size_t>(n), -1);
    std::vector<int32_t> permOf(static_cast<size_t>(n), -1);
    pf->diag.assign(static_cast<size_t>(n), 0.0);
    pf->sub_.assign(static_cast<size_t>(n), 0.0);
    pf->rowsStart.assign(S + 1, 0);
    int32_t next = 0;
    for (int32_t s = 0; s < S; ++s)
    {
        d->first[s] = next;
        d->nrow[s] = pf->order[s] - pf->eliminated[s];
        for (int32_t k = 0; k < pf->eliminated[s]; ++k)
        {
            const int32_t eq = frontRows[s][k];
            if (eq < 0 || eq >= n || pf->pivotOf[eq] >= 0)
                fail(OOSG_INTERNAL, "factor: an equation was eliminated twice");
            pf->pivotOf[eq] = next;
            permOf[next] = eq;
            pf->diag[next] = dOf[s][k];
            pf->sub_[next] = sOf[s][k];
            ++next;
        }
        pf->rowsStart[s + 1] = pf->rowsStart[s] + pf->order[s];
    }
    if (next != n)
        fail(OOSG_INTERNAL, "factor: " + std::to_string(n - next) + " equations were not eliminated");
    pf->panelRows.reserve(static_cast<size_t>(pf->rowsStart[S]));
    for (int32_t s = 0; s < S; ++s)
        pf->panelRows.insert(pf->panelRows.end(), frontRows[s].begin(), frontRows[s].end());
    d->rowStart.assign(S + 1, 0);
    for (int32_t s = 0; s < S; ++s)
        d->rowStart[s + 1] = d->rowStart[s] + d->nrow[s];
    d->rows.resize(static_cast<size_t>(d->rowStart[S]));
    d->rel.assign(d->rows.size(), 0);
    for (int32_t s = 0; s < S; ++s)
        for (int32_t i = 0; i < d->nrow[s]; ++i)
            d->rows[d->rowStart[s] + i] = pf->pivotOf[frontRows[s][pf->eliminated[s] + i]];
    {
        // Each child's rows → positions in its parent's panel rows.
        std::vector<int32_t> posOf(static_cast<size_t>(n), -1);
        for (int32_t s = 0; s < S; ++s)
        {
            if (p.childStart[s] == p.childStart[s + 1])
                continue;
            for (int32_t k = 0; k < pf->order[s]; ++k)
                posOf[frontRows[s][k]] = k;
            for (int32_t x = p.childStart[s]; x < p.childStart[s + 1]; ++x)
            {
                const int32_t ch = p.children[x];
                for (int32_t i = 0; i < d->nrow[ch]; ++i)
                {
                    const int32_t at = posOf[frontRows[ch][pf->eliminated[ch] + i]];
                    if (at < 0)
                        fail(OOSG_INTERNAL, "factor: a row of a child is not in its parent's front");
                    d->rel[d->rowStart[ch] + i] = at;
                }
            }
            for (int32_t k = 0; k < pf->order[s]; ++k)
                posOf[frontRows[s][k]] = -1;
        }
    }
    for (int32_t s = 0; s < S; ++s)
        if (d->parent[s] < 0 && d->nrow[s] > 0)
            fail(OOSG_INTERNAL, "factor: a root front kept rows");
    deriveLevels(*d);
    uploadSweepPlan(*d);
    // The inverses of the unit lower triangles: the diagonal blocks, then the wide triangles' whole inverses.
    allocateCounted(p, f.linv, static_cast<size_t>(std::max<int64_t>(d->linvOff[S], 1)) * sizeof(double),
                    "factor: inverses");
    // The sweeps' vectors of the pivoted structure, counted in the plan's budget.
    ensureCounted(p, d->solveX, static_cast<size_t>(std::max(n, 1)) * 16 * sizeof(double), "factor: solve vectors");
    ensureCounted(p, d->solveU, static_cast<size_t>(std::max<int64_t>(d->rowStart[S], 1)) * 16 * sizeof(double),
                  "factor: solve vectors");
    ensureCounted(p, d->solvePartials, static_cast<size_t>(std::max<int64_t>(d->partialSlots, 1)) * 64 * 16 *
                                           sizeof(double),
                  "factor: solve vectors");
    std::vector<int2> blockItems;
    for (int32_t s = 0; s < S; ++s)
        for (int32_t J = 0; J * oosg::PANEL_BLOCK < d->ncol[s]; ++J)
            blockItems.push_back(make_int2(s, J));
    uploadInto(p.pivotInverseItems, blockItems, st, "factor: inverse blocks");
    check(oosg::launchUnitBlockInverses(d->dp, f.panels.as<double>(), f.linv.as<double>(),
                                        p.pivotInverseItems.as<int2>(), static_cast<int>(blockItems.size()), st),
          "factor: inverses");
    for (int32_t l = 0; l < d->levels; ++l)
        for (const oosg_plan::InverseRound& round : d->inverseRounds[l])
        {
            check(oosg::launchTriangleInverse<double>(d->dp, f.panels.as<double>(), f.linv.as<double>(),
                                                      d->dInverseItems.as<int2>() + round.rowOffset, round.rowCount,
                                                      -1, st),
                  "factor: triangle inverses");
            check(oosg::launchTriangleInverse<double>(d->dp, f.panels.as<double>(), f.linv.as<double>(),
                                                      d->dInverseItems.as<int2>() + round.offset, round.count,
                                                      round.step, st),
                  "factor: triangle inverses");
        }
    f.dDiag = upload(pf->diag, st, "factor: pivots");
    pf->sub = upload(pf->sub_, st, "factor: pivots");
    pf->perm = upload(permOf, st, "factor: pivot order");
    f.diag.assign(static_cast<size_t>(n), 0.0);
    for (int32_t eq = 0; eq < n; ++eq)
        f.diag[eq] = pf->diag[pf->pivotOf[eq]];
    check(hipStreamSynchronize(st), "factor");
    pf->structure = std::move(d);
    f.pivoted = std::move(pf);
}

// X (row-major n × k, k <= 16, in p.solveX) ← factor⁻¹·X. sub: the subdiagonal of a block diagonal D (null: D is
// diagonal).
// The forward sweep of the levels [l0, l1) on p.solveX, on `st`.
template <typename TP>
void forwardSweep(oosg_plan& p, const TP* panels, const TP* linv, int k, int32_t l0, int32_t l1, hipStream_t st)
{
    double* X = p.solveX.as<double>();
    double* U = p.solveU.as<double>();
    const int32_t* tiny = p.dTinyList.as<int32_t>();
    const int32_t* others = p.dOtherList.as<int32_t>();
    const int4* forwardTiles = p.dForwardTiles.as<int4>();
    const int4* subtractions = p.dForwardSubtractions.as<int4>();
    const int4* triangles = p.dForwardTriangles.as<int4>();
    const int4* triangleSums = p.dForwardTriangleSums.as<int4>();
    double* partials = p.solvePartials.as<double>();
    if (p.partialSlots > 0 && partials == nullptr)
        fail(OOSG_INTERNAL, "solve: the sweeps' partial sums have no buffer");
    // The profile counts bytes as the work of the sweeps (they are bound by HBM): the values of L (or of an inverse)
    // a launch reads once, and the right-hand-side rows it reads or writes once.
    const double T = sizeof(TP), row = 8.0 * k;
    const auto timed = [&](int kind, double work, auto&& launch)
    {
        const int slot = p.profiler.begin(kind, work, st);
        launch();
        p.profiler.end(slot, st);
    };
    for (int32_t l = l0; l < l1; ++l)
    {
        const oosg_plan::SweepWork& w = p.sweepWork[l];
        // Gathers and triangles, the wide supernodes' triangles (products with their whole inverses in chunks of
        // blocks, whose sums replace the rows of X once every product has read X), then every L_RF row tile.
        timed(PROFILE_FORWARD, w.triangleValues * T + w.forwardRows * row, [&] {
            check(oosg::launchTinyForward<TP>(p.dp, panels, linv, X, U, k, tiny + p.tinyStart[l],
                                              p.tinyStart[l + 1] - p.tinyStart[l], st),
                  "solve: forward sweep (tiny supernodes)");
            const int count = p.otherStart[l + 1] - p.otherStart[l];
            check(oosg::launchGather(p.dGatherChunks.as<oosg::GatherChunk>() + p.gatherChunkStart[l],
                                     p.gatherChunkStart[l + 1] - p.gatherChunkStart[l], p.dGatherStart.as<int32_t>(),
                                     p.dGatherRows.as<int32_t>(), X, U, k, st),
                  "solve: forward sweep (gathers)");
            check(oosg::launchMediumForward<TP>(p.dp, panels, linv, X, U, k,
                                                p.dMediumList.as<int32_t>() + p.mediumStart[l],
                                                p.mediumStart[l + 1] - p.mediumStart[l], st),
                  "solve: forward sweep (medium supernodes)");
            check(oosg::launchForward<TP>(p.dp, panels, linv, X, U, k, others + p.otherStart[l], count, nullptr, 0,
                                          partials, st, true),
                  "solve: forward sweep");
        });
        const int items = static_cast<int>(p.forwardTriangleStart[l + 1] - p.forwardTriangleStart[l]);
        const int sums = static_cast<int>(p.triangleSumStart[l + 1] - p.triangleSumStart[l]);
        if (items > 0)
        {
            timed(PROFILE_TRIANGLE_PRODUCTS, w.wideValues * T + w.wideRows * row, [&] {
                check(oosg::launchTriangleProducts<TP>(p.dp, linv, X, k, triangles + p.forwardTriangleStart[l], items,
                                                       0, partials, st),
                      "solve: forward sweep (triangles)");
            });
            check(oosg::launchSubtractPartials(X, nullptr, k, triangleSums + p.triangleSumStart[l], sums, partials, st),
                  "solve: forward sweep (triangles)");
        }
        const int tiles = static_cast<int>(p.forwardTileStart[l + 1] - p.forwardTileStart[l]);
        const int subtracted = static_cast<int>(p.forwardSubtractionStart[l + 1] - p.forwardSubtractionStart[l]);
        if (tiles > 0)
            timed(PROFILE_FORWARD_ROWS, w.rowValues * T + w.rowTileRows * row, [&] {
                check(oosg::launchForward<TP>(p.dp, panels, linv, X, U, k, nullptr, 0,
                                              forwardTiles + p.forwardTileStart[l], tiles, partials, st),
                      "solve: forward sweep (rows)");
                check(oosg::launchSubtractPartials(X, U, k, subtractions + p.forwardSubtractionStart[l], subtracted,
                                                   partials, st),
                      "solve: forward sweep (rows)");
            });
    }
}

// p.solveX ← L⁻ᵀ·p.solveX: the backward sweep, levels from the top (rows outside the levels, a partial factor's
// trailing ones, are read and kept).
template <typename TP>
void backwardSweep(oosg_plan& p, const TP* panels, const TP* linv, int k, int32_t l0 = 0, int32_t l1 = -1)
{
    if (l1 < 0)
        l1 = p.levels;
    const hipStream_t st = p.stream;
    double* X = p.solveX.as<double>();
    const int32_t* tiny = p.dTinyList.as<int32_t>();
    const int32_t* others = p.dOtherList.as<int32_t>();
    const int4* backwardBlocks = p.dBackwardBlocks.as<int4>();
    const int4* subtractions = p.dBackwardSubtractions.as<int4>();
    double* partials = p.solvePartials.as<double>();
    if (p.partialSlots > 0 && partials == nullptr)
        fail(OOSG_INTERNAL, "solve: the sweeps' partial sums have no buffer");
    const int4* triangles = p.dBackwardTriangles.as<int4>();
    const int4* triangleSums = p.dBackwardTriangleSums.as<int4>();
    const double T = sizeof(TP), row = 8.0 * k;
    const auto timed = [&](int kind, double work, auto&& launch)
    {
        const int slot = p.profiler.begin(kind, work, st);
        launch();
        p.profiler.end(slot, st);
    };
    for (int32_t l = l1 - 1; l >= l0; --l)
    {
        const oosg_plan::SweepWork& w = p.sweepWork[l];
        // The L_RFᵀ products of every supernode first (column blocks), then the wide supernodes' triangles (with
        // their whole inverses), then the other triangles.
        const int blocks = static_cast<int>(p.backwardBlockStart[l + 1] - p.backwardBlockStart[l]);
        const int subtracted = static_cast<int>(p.backwardSubtractionStart[l + 1] - p.backwardSubtractionStart[l]);
        if (blocks > 0)
            timed(PROFILE_BACKWARD_COLUMNS, w.rowValues * T + w.columnRows * row, [&] {
                check(oosg::launchBackward<TP>(p.dp, panels, linv, X, k, others, 0,
                                               backwardBlocks + p.backwardBlockStart[l], blocks, partials, st),
                      "solve: backward sweep (columns)");
                check(oosg::launchSubtractPartials(X, nullptr, k, subtractions + p.backwardSubtractionStart[l],
                                                   subtracted, partials, st),
                      "solve: backward sweep (columns)");
            });
        const int items = static_cast<int>(p.backwardTriangleStart[l + 1] - p.backwardTriangleStart[l]);
        const int sums = static_cast<int>(p.triangleSumStart[l + 1] - p.triangleSumStart[l]);
        if (items > 0)
        {
            timed(PROFILE_TRIANGLE_PRODUCTS, w.wideValues * T + w.wideRows * row, [&] {
                check(oosg::launchTriangleProducts<TP>(p.dp, linv, X, k, triangles + p.backwardTriangleStart[l],
                                                       items, 1, partials, st),
                      "solve: backward sweep (triangles)");
            });
            check(oosg::launchSubtractPartials(X, nullptr, k, triangleSums + p.triangleSumStart[l], sums, partials, st),
                  "solve: backward sweep (triangles)");
        }
        timed(PROFILE_BACKWARD, w.triangleValues * T + w.backwardRows * row, [&] {
            check(oosg::launchTinyBackward<TP>(p.dp, panels, linv, X, k, tiny + p.tinyStart[l],
                                               p.tinyStart[l + 1] - p.tinyStart[l], st),
                  "solve: backward sweep (tiny supernodes)");
            check(oosg::launchMediumBackward<TP>(p.dp, panels, linv, X, k, p.dMediumList.as<int32_t>() + p.mediumStart[l],
                                                 p.mediumStart[l + 1] - p.mediumStart[l], st),
                  "solve: backward sweep (medium supernodes)");
            check(oosg::launchBackward<TP>(p.dp, panels, linv, X, k, others + p.otherStart[l],
                                           p.otherStart[l + 1] - p.otherStart[l], backwardBlocks, 0, partials, st),
                  "solve: backward sweep");
        });
    }
}

// p.solveX ← factor⁻¹·p.solveX: the forward sweep from level `fromLevel` (the levels below it already swept), D⁻¹,
// the backward sweep.
template <typename TP>
void runSweeps(oosg_plan& p, const TP* panels, const TP* linv, const double* diag, int k, const double* sub = nullptr,
               int32_t fromLevel = 0)
{
    const hipStream_t st = p.stream;
    double* X = p.solveX.as<double>();
    forwardSweep<TP>(p, panels, linv, k, fromLevel, p.levels, st);
    if (sub != nullptr)
        check(oosg::launchBlockDiagonalSolve(X, diag, sub, p.n, k, st), "solve: block diagonal");
    else
        check(oosg::launchDiagonalSolve(X, diag, p.n, k, st), "solve: diagonal");
    backwardSweep<TP>(p, panels, linv, k);
}

// An out-of-core factor's solve: p.solveX ← factor⁻¹·p.solveX with each level's panels and inverses copied from the
// spool into window l mod 2 on the copy stream, the next level's copy beside this level's sweep; after the forward
// sweep the windows hold the two top levels, so the backward sweep starts without a copy.
template <typename TP> void runSweepsOutOfCore(const oosg_factor& f, int k)
{
    oosg_plan& p = *f.plan;
    const hipStream_t st = p.stream;
    TP* panels = f.panels.as<TP>();
    TP* linv = f.linv.as<TP>();
    const TP* spool = static_cast<const TP*>(f.spool);
    const int32_t L = p.levels;
    const auto load = [&](int32_t level)
    {
        const int w = level % 2;
        check(hipStreamWaitEvent(p.copyStream, p.oocFree[w], 0), "solve: spool");
        const int64_t values = p.spoolPanelBase[level + 1] - p.spoolPanelBase[level];
        const int64_t inverses = p.spoolLinvBase[level + 1] - p.spoolLinvBase[level];
        if (values > 0)
            check(hipMemcpyAsync(panels + w * p.windowPanels, spool + p.spoolPanelBase[level],
                                 static_cast<size_t>(values) * sizeof(TP), hipMemcpyHostToDevice, p.copyStream),
                  "solve: spool");
        if (inverses > 0)
            check(hipMemcpyAsync(linv + w * p.windowLinv, spool + p.spoolPanelBase[L] + p.spoolLinvBase[level],
                                 static_cast<size_t>(inverses) * sizeof(TP), hipMemcpyHostToDevice, p.copyStream),
                  "solve: spool");
        check(hipEventRecord(p.oocReady[w], p.copyStream), "solve: spool");
        f.traffic.read_bytes += static_cast<uint64_t>(values + inverses) * sizeof(TP);
        f.traffic.read_requests += (values > 0 ? 1 : 0) + (inverses > 0 ? 1 : 0);
    };
    const auto sweep = [&](int32_t level, bool forward)
    {
        const int w = level % 2;
        check(hipStreamWaitEvent(st, p.oocReady[w], 0), "solve: spool");
        if (forward)
            forwardSweep<TP>(p, panels + w * p.windowPanels, linv + w * p.windowLinv, k, level, level + 1, st);
        else
            backwardSweep<TP>(p, panels + w * p.windowPanels, linv + w * p.windowLinv, k, level, level + 1);
        check(hipEventRecord(p.oocFree[w], st), "solve: spool");
    };
    if (L > 0)
        load(0);
    for (int32_t l = 0; l < L; ++l)
    {
        if (l + 1 < L)
            load(l + 1);
        sweep(l, true);
    }
    check(oosg::launchDiagonalSolve(p.solveX.as<double>(), f.dDiag.as<double>(), p.n, k, st), "solve: diagonal");
    for (int32_t l = L - 1; l >= 0; --l)
    {
        if (l - 1 >= 0 && l - 1 <= L - 3)
            load(l - 1); // into the window of level l + 1, swept already
        sweep(l, false);
    }
}

// p.solveX ← factor⁻¹·p.solveX (row-major n × k), as a graph replay unless the profiler is on; with fromLevel > 0
// the forward sweep of the levels below it is done already (a combined factorisation and solve), and the rest runs
// launch by launch.
void sweeps(const oosg_factor& f, int k, int32_t fromLevel = 0)
{
    oosg_plan& p = *f.plan;
    if (p.trailing >= 0)
        fail(OOSG_INVALID_ARGUMENT, "solve: a partial factor solves by oosg_partial_condense_loads and "
                                    "oosg_partial_recover (its trailing equations are not eliminated)");
    if (p.outOfCore)
    {
        if (f.precision == OOSG_FP32)
            runSweepsOutOfCore<float>(f, k);
        else
            runSweepsOutOfCore<double>(f, k);
        return;
    }
    if (fromLevel > 0)
    {
        if (f.precision == OOSG_FP32)
            runSweeps<float>(p, f.panels.as<float>(), f.linv.as<float>(), f.dDiag.as<double>(), k, nullptr, fromLevel);
        else
            runSweeps<double>(p, f.panels.as<double>(), f.linv.as<double>(), f.dDiag.as<double>(), k, nullptr,
                              fromLevel);
        return;
    }
    if (f.precision == OOSG_MIXED)
    {
        // FP32 levels below the cut, FP64 above: forward up, D⁻¹, backward down.
        const int32_t cut = f.fp32Levels;
        const auto body = [&]
        {
            forwardSweep<float>(p, f.panels32.as<float>(), f.linv32.as<float>(), k, 0, cut, p.stream);
            forwardSweep<double>(p, f.panels.as<double>(), f.linv.as<double>(), k, cut, p.levels, p.stream);
            check(oosg::launchDiagonalSolve(p.solveX.as<double>(), f.dDiag.as<double>(), p.n, k, p.stream),
                  "solve: diagonal");
            backwardSweep<double>(p, f.panels.as<double>(), f.linv.as<double>(), k, cut, p.levels);
            backwardSweep<float>(p, f.panels32.as<float>(), f.linv32.as<float>(), k, 0, cut);
        };
        if (!p.graphs || p.profiler.enabled)
            body();
        else
            p.sweepGraph.run(p.stream,
                             {f.panels.ptr, f.linv.ptr, f.panels32.ptr, f.linv32.ptr, f.dDiag.ptr, p.solveX.ptr,
                              p.solveU.ptr, p.solvePartials.ptr,
                              reinterpret_cast<const void*>(static_cast<intptr_t>(k * 4096 + 2048 + cut))},
                             body);
        return;
    }
    if (f.pivoted)
    {
        // In pivot order: X permuted into the structure's own X, swept, and permuted back.
        oosg_plan& d = *f.pivoted->structure;
        const int64_t n = p.n;
        const int32_t* perm = f.pivoted->perm.as<int32_t>();
        check(oosg::launchPermuteRows(p.solveX.as<double>(), d.solveX.as<double>(), perm, n, k, false, p.stream),
              "solve: pivot order");
        runSweeps<double>(d, f.panels.as<double>(), f.linv.as<double>(), f.dDiag.as<double>(), k,
                          f.pivoted->sub.as<double>());
        check(oosg::launchPermuteRows(d.solveX.as<double>(), p.solveX.as<double>(), perm, n, k, true, p.stream),
              "solve: pivot order");
        return;
    }
    const auto body = [&]
    {
        if (f.precision == OOSG_FP32)
            runSweeps<float>(p, f.panels.as<float>(), f.linv.as<float>(), f.dDiag.as<double>(), k);
        else
            runSweeps<double>(p, f.panels.as<double>(), f.linv.as<double>(), f.dDiag.as<double>(), k);
    };
    if (!p.graphs || p.profiler.enabled)
    {
        body();
        return;
    }
    p.sweepGraph.run(p.stream,
                     {f.panels.ptr, f.linv.ptr, f.dDiag.ptr, p.solveX.ptr, p.solveU.ptr,
                      p.solvePartials.ptr, reinterpret_cast<const void*>(static_cast<intptr_t>(k * 2 + f.precision))},
                     body);
}

// b (device, column-major, ld) ← factor⁻¹·b, by blocks of 16 columns.
void solveDevice(const oosg_factor& f, double* b, int64_t ld, int64_t nrhs)
{
    oosg_plan& p = *f.plan;
    const int64_t n = p.n;
    const int64_t width = std::min<int64_t>(nrhs, 16);
    (void)width;
    // Sized for 16 columns whatever the call: the buffers (and the graphs replaying the sweeps on them) stay put.
    ensureSolveBuffers(p);
    const double* scale = scaledPrecision(f.precision) ? f.scale.as<double>() : nullptr;
    for (int64_t c0 = 0; c0 < nrhs; c0 += 16)
    {
        if (c0 > 0)
            pollCancellation(p);
        const int k = static_cast<int>(std::min<int64_t>(16, nrhs - c0));
        double* block = b + c0 * ld;
        check(oosg::launchToRowMajor(block, ld, p.solveX.as<double>(), n, k, scale, p.stream), "solve: layout");
        sweeps(f, k);
        check(oosg::launchFromRowMajor(p.solveX.as<double>(), block, ld, n, k, scale, 0, p.stream), "solve: layout");
    }
}

// The right-hand sides b (column-major, leading dimension ld, k <= 16 columns) into the plan's cgB (row-major) and
// X1 = S·b (row-major; scale null: a copy), their column maxima into slot 0 of the plan's norms: one pass, on the
// plan's stream.
void prepareRightHandSides(oosg_plan& p, const double* b, int64_t ld, double* X1, const double* scale, int k)
{
    unsigned long long* norm = p.solveNorms.as<unsigned long long>();
    check(oosg::launchFillInt32(reinterpret_cast<int32_t*>(norm), 32, 0, p.stream), "solve: norms");
    check(oosg::launchRightHandSides(b, ld, p.cgB.as<double>(), X1, scale, p.n, k, norm, p.stream), "solve: layout");
}

// Column maxima |v| of row-major n × k vectors into slot `slot` (of 3) of the plan's norms, on the stream; with a
// row scaling S, of S·v (divide: S⁻¹·v).
void startColumnMax(oosg_plan& p, const double* v, int k, int slot, const double* scale = nullptr, int divide = 0)
{
    unsigned long long* out = p.solveNorms.as<unsigned long long>() + 16 * slot;
    check(oosg::launchFillInt32(reinterpret_cast<int32_t*>(out), 32, 0, p.stream), "solve: norms");
    check(oosg::launchColumnMaxRows(v, p.n, k, out, p.stream, scale, divide), "solve: norms");
}

// The first `slots` slots of column maxima, read back with one copy and one wait.
std::vector<std::vector<double>> finishColumnMax(oosg_plan& p, int slots, int k)
{
    std::vector<unsigned long long> bits(static_cast<size_t>(16 * slots));
    check(hipMemcpyAsync(bits.data(), p.solveNorms.ptr, bits.size() * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, p.stream),
          "solve: norms");
    check(hipStreamSynchronize(p.stream), "solve: norms");
    std::vector<std::vector<double>> out(static_cast<size_t>(slots), std::vector<double>(static_cast<size_t>(k)));
    for (int q = 0; q < slots; ++q)
        for (int h = 0; h < k; ++h)
            std::memcpy(&out[q][h], &bits[16 * q + h], sizeof(double));
    return out;
}

// The solve buffers of a plan, sized for 16 right-hand sides.
// (Reserved with the plan: nothing is allocated here unless a buffer was given back.)
void ensureSolveBuffers(oosg_plan& p)
{
    reserveSolveBuffers(p);
}

// The residual product of a step: its bytes for the profile.
double residualBytes(const oosg_matrix& m, int64_t n, int k)
{
    return static_cast<double>(m.csrFromCsc.size()) * 12.0 + static_cast<double>(n + 1) * 8.0 +
           2.0 * static_cast<double>(n) * k * 8.0;
}

// ---- block conjugate gradients ------------------------------------------------------------------------------------

// The eigenvalues w and eigenvectors (columns of v, row-major k × k) of a symmetric k × k matrix `a` (row-major), by
// cyclic Jacobi rotations (k <= 16: a few sweeps).
void symmetricEigen(int k, std::vector<double> a, std::vector<double>& w, std::vector<double>& v)
{
    v.assign(static_cast<size_t>(k * k), 0.0);
    for (int i = 0; i < k; ++i)
        v[i * k + i] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep)
    {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j)
                (i == j ? diag : off) += a[i * k + j] * a[i * k + j];
        if (off <= 1e-32 * diag || off == 0.0)
            break;
        for (int p = 0; p < k - 1; ++p)
            for (int q = p + 1; q < k; ++q)
            {
                const double apq = a[p * k + q];
                if (apq == 0.0)
                    continue;
                const double theta = (a[q * k + q] - a[p * k + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int r = 0; r < k; ++r)
                {
                    const double arp = a[r * k + p], arq = a[r * k + q];
                    a[r * k + p] = c * arp - sn * arq;
                    a[r * k + q] = sn * arp + c * arq;
                }
                for (int r = 0; r < k; ++r)
                {
                    const double apr = a[p * k + r], aqr = a[q * k + r];
                    a[p * k + r] = c * apr - sn * aqr;
                    a[q * k + r] = sn * apr + c * aqr;
                }
                for (int r = 0; r < k; ++r)
                {
                    const double vrp = v[r * k + p], vrq = v[r * k + q];
                    v[r * k + p] = c * vrp - sn * vrq;
                    v[r * k + q] = sn * vrp + c * vrq;
                }
            }
    }
    w.resize(static_cast<size_t>(k));
    for (int i = 0; i < k; ++i)
        w[i] = a[i * k + i];
}

// T (16 × 16, row a, column j) such that Tᵀ·G·T is the identity on the directions kept and T's other columns are zero,
// for the Gram matrix G = Wᵀ·A·W of a block of directions W (16 × 16 with k used): G scaled to a unit diagonal,
// its eigenvectors with eigenvalues below 1e-12 of the largest (linearly dependent directions, or zero ones)
// dropped.
std::vector<double> orthonormalizer(const double* G, int k)
{
    std::vector<double> d(static_cast<size_t>(k)), g(static_cast<size_t>(k * k)), w, v;
    for (int a = 0; a < k; ++a)
        d[a] = G[a * 16 + a] > 0.0 && std::isfinite(G[a * 16 + a]) ? 1.0 / std::sqrt(G[a * 16 + a]) : 0.0;
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b)
            g[a * k + b] = d[a] * 0.5 * (G[a * 16 + b] + G[b * 16 + a]) * d[b];
    symmetricEigen(k, g, w, v);
    double largest = 0.0;
    for (const double x : w)
        largest = std::max(largest, x);
    std::vector<double> T(256, 0.0);
    for (int j = 0; j < k; ++j)
    {
        if (!(w[j] > 1e-12 * largest))
            continue;
        const double s = 1.0 / std::sqrt(w[j]);
        for (int a = 0; a < k; ++a)
            T[a * 16 + j] = d[a] * v[a * k + j] * s;
    }
    return T;
}

// Block conjugate gradients on A (symmetric positive definite) preconditioned by the factor, over the k columns at
// once (O'Leary, "The block conjugate gradient algorithm", Linear Algebra Appl. 29, 1980), run on the scaled system
// A_s·y = S·b, A_s = S·A·S, x = S·y, whose factor the FP32 fronts hold (S: powers of two, so every scaling is exact
// and no vector is rescaled between the steps). The search block P is kept A_s-orthonormal (Pᵀ·A_s·P = I), so
// α = Pᵀ·R and β = −(A_s·P)ᵀ·Z. With Z = A_s⁻¹·R (the sweeps, in place on the sweep vectors):
//   W = Z − P·((A_s·P)ᵀ·Z),  G = Wᵀ·A_s·W,  T with Tᵀ·G·T = I (dependent directions dropped),
//   P = W·T, A_s·P = (A_s·W)·T,  α = Pᵀ·R,  Y += P·α,  R −= (A_s·P)·α,
// the last four and the copy of R into the sweep vectors in one pass (launchBlockStep), which also takes the column
// maxima of the unscaled residual S⁻¹·R and solution S·Y. The k columns share one Krylov space: the few directions
// where the FP32 factor is least accurate (the smooth modes of a thin plate) serve every right-hand side. Stopping
// rule: the recurrence's backward error first; once it meets the target, the true residual S·(b − A·x) decides, and
// replaces the recurrence's when it does not. Buffers: the sweep vectors hold W, P holds A_s·W before it holds P,
// Q holds A_s·P.
void solveBlockPcg(const oosg_factor& f, const oosg_matrix& m, const double* b, double* x, int64_t ld, int k,
                   double target, int32_t maxSteps, oosg_solve_report& report, int32_t aheadLevel = 0)
{
    oosg_plan& p = *f.plan;
    const hipStream_t st = p.stream;
    const int64_t n = p.n;
    double* B = p.cgB.as<double>();
    double* X = p.cgX.as<double>();
    double* R = p.cgR.as<double>();
    double* Z = p.solveX.as<double>();
    double* P = p.cgP.as<double>();
    double* Q = p.cgQ.as<double>();
    double* grams = p.cgDots.as<double>();
    double* coef = grams + 512; // T and U = T·α, after the two Gram products
    double* partial = p.cgPartial.as<double>();
    unsigned long long* norms = p.solveNorms.as<unsigned long long>();
    const double* scale = scaledPrecision(f.precision) ? f.scale.as<double>() : nullptr;
    const int64_t* offsets = m.dCsrOffsets.as<int64_t>();
    const int32_t* cols = m.dCsrCols.as<int32_t>();
    const double* values = m.dCsrValues.as<double>();
    const int32_t* groups = m.dCsrGroups.as<int32_t>();
    const int64_t groupCount = static_cast<int64_t>(m.csrGroups.size()) - 1;
    const size_t vectorBytes = static_cast<size_t>(n) * k * sizeof(double);
    const auto product = [&](const double* in, const double* minus, double* out)
    {
        const int slot = p.profiler.begin(PROFILE_RESIDUAL, residualBytes(m, n, k), st);
        check(oosg::launchMultiplyRows(offsets, cols, values, groups, groupCount, in, minus, out, k, st, scale),
              "solve: product");
        p.profiler.end(slot, st);
    };
    const auto gram = [&](std::initializer_list<std::pair<const double*, const double*>> pairs)
    {
        oosg::GramPairs g{};
        g.pairs = 0;
        for (const auto& pr : pairs)
        {
            g.u[g.pairs] = pr.first;
            g.v[g.pairs] = pr.second;
            ++g.pairs;
        }
        check(oosg::launchBlockGram(g, n, k, partial, grams, st), "solve: Gram products");
    };

    // With aheadLevel > 0, B holds b, the sweep vectors S·b, swept forward below that level, and the first slot of
    // the norms ‖b‖∞ by column.
    if (aheadLevel == 0)
        prepareRightHandSides(p, b, ld, Z, scale, k);
    const std::vector<double> bNorm = finishColumnMax(p, 1, k)[0];
    check(oosg::launchFillDouble(X, static_cast<int64_t>(n) * k, static_cast<int64_t>(n) * k, 1, 0.0, st),
          "solve: zero");
    check(oosg::launchScaleRows(B, R, scale, n, k, st), "solve: residual");
    // The backward errors from the norms of slots 1 (the residual) and 2 (the solution).
    const auto backwardErrors = [&]()
    {
        const std::vector<std::vector<double>> norms3 = finishColumnMax(p, 3, k);
        std::vector<double> e(static_cast<size_t>(k));
        for (int h = 0; h < k; ++h)
        {
            const double denominator = m.normInf * norms3[2][h] + bNorm[h];
            const double v = denominator > 0 ? norms3[1][h] / denominator : norms3[1][h];
            e[h] = v != v ? INFINITY : v;
        }
        return e;
    };
    std::vector<double> eta(static_cast<size_t>(k), INFINITY);
    int32_t solves = 0;
    bool converged = false;
    for (int32_t iteration = 0;; ++iteration)
    {
        // Z = A_s⁻¹·R, then W = Z − P·((A_s·P)ᵀ·Z) in Z.
        sweeps(f, k, iteration == 0 ? aheadLevel : 0);
        ++solves;
        if (iteration > 0)
        {
            gram({{Q, Z}});
            oosg::BlockTerm t{};
            t.u = P;
            t.deviceC = grams;
            t.scale = -1.0;
            t.c0 = 1.0;
            check(oosg::launchBlockCombine(Z, t, n, k, st), "solve: directions");
        }
        // A_s·W (in P), G = Wᵀ·A_s·W and F = Wᵀ·R, read back.
        product(Z, nullptr, P);
        gram({{Z, P}, {Z, R}});
        std::vector<double> G(512);
        check(hipMemcpyAsync(G.data(), grams, G.size() * sizeof(double), hipMemcpyDeviceToHost, st),
              "solve: Gram products");
        check(hipStreamSynchronize(st), "solve: Gram products");
        // T, then U = T·α with α = Tᵀ·F (16 × 16, row j, column h): U = T·Tᵀ·F.
        std::vector<double> coefficients(512, 0.0);
        const std::vector<double> T = orthonormalizer(G.data(), k);
        std::copy(T.begin(), T.end(), coefficients.begin());
        std::vector<double> alpha(256, 0.0);
        for (int j = 0; j < k; ++j)
            for (int h = 0; h < k; ++h)
            {
                double sum = 0.0;
                for (int a = 0; a < k; ++a)
                    sum += T[a * 16 + j] * G[256 + a * 16 + h];
                alpha[j * 16 + h] = sum;
            }
        for (int a = 0; a < k; ++a)
            for (int h = 0; h < k; ++h)
            {
                double sum = 0.0;
                for (int j = 0; j < k; ++j)
                    sum += T[a * 16 + j] * alpha[j * 16 + h];
                coefficients[256 + a * 16 + h] = sum;
            }
        check(hipMemcpyAsync(coef, coefficients.data(), coefficients.size() * sizeof(double), hipMemcpyHostToDevice,
                             st),
              "solve: coefficients");
        // Q = (A_s·W)·T, P = W·T, Y += W·U, R −= (A_s·W)·U, Z = R, and the norms of slots 1 and 2.
        check(oosg::launchFillInt32(reinterpret_cast<int32_t*>(norms + 16), 64, 0, st), "solve: norms");
        check(oosg::launchBlockStep(Z, P, Q, X, R, coef, scale, n, k, norms + 16, st), "solve: step");
        const std::vector<double> estimate = backwardErrors();
        bool met = true;
        for (int h = 0; h < k; ++h)
            met = met && estimate[h] <= target;
        if (met || solves >= maxSteps)
        {
            // The true residual S·(b − A·x), into R.
            product(X, B, R);
            startColumnMax(p, R, k, 1, scale, 1);
            startColumnMax(p, X, k, 2, scale, 0);
            const std::vector<double> trueEta = backwardErrors();
            converged = true;
            for (int h = 0; h < k; ++h)
            {
                eta[h] = trueEta[h];
                converged = converged && trueEta[h] <= target;
            }
            if (converged || solves >= maxSteps)
                break;
            // Residual replacement: the recurrence drifted from b − A·x; the sweeps start from the true residual.
            check(hipMemcpyAsync(Z, R, vectorBytes, hipMemcpyDeviceToDevice, st), "solve: residual");
        }
        pollCancellation(p);
    }
    check(oosg::launchFromRowMajor(X, x, ld, n, k, scale, 0, st), "solve: layout");
    report.steps = solves;
    report.converged = converged ? 1 : 0;
    report.backward_error = 0.0;
    for (int h = 0; h < k; ++h)
        report.backward_error = std::max(report.backward_error, eta[h]);
}

// Iterative refinement in double precision on row-major vectors: x_0 = 0, r_0 = b; d = factor⁻¹·r, x += d,
// r = b − A·x; stops at the backward-error target, after max_steps solves, or when a step does not halve the
// backward error (the factor is then not accurate enough for the matrix). The first solution stays in the sweeps'
// vectors (no copy) unless a second step needs them.
void solveRefinement(const oosg_factor& f, const oosg_matrix& m, const double* b, double* x, int64_t ld, int k,
                     double target, int32_t maxSteps, oosg_solve_report& report, int32_t aheadLevel = 0)
{
    oosg_plan& p = *f.plan;
    const hipStream_t st = p.stream;
    const int64_t n = p.n;
    double* B = p.cgB.as<double>();
    double* R = p.cgR.as<double>();
    double* sweepVectors = p.solveX.as<double>();
    double* X = sweepVectors;
    const double* scale = scaledPrecision(f.precision) ? f.scale.as<double>() : nullptr;
    // B holds b, the sweep vectors S·b (swept forward below aheadLevel when it is > 0) and the first slot of the norms
    // ‖b‖∞ by column (read with the first step's norms).
    if (aheadLevel == 0)
        prepareRightHandSides(p, b, ld, sweepVectors, scale, k);
    std::vector<double> bNorm;
    bool written = false; // x holds the first solution already
    double previous = INFINITY;
    report.converged = 0;
    for (int32_t step = 1; step <= std::max(1, maxSteps); ++step)
    {
        // x_0 = 0, r_0 = b: the first correction is the first solution.
        if (step == 1)
        {
            sweeps(f, k, aheadLevel);
            if (scale != nullptr)
                check(oosg::launchScaleRows(sweepVectors, sweepVectors, scale, n, k, st), "solve: scaling");
        }
        else
        {
            if (X == sweepVectors)
            {
                X = p.cgX.as<double>();
                check(hipMemcpyAsync(X, sweepVectors, static_cast<size_t>(n) * k * sizeof(double),
                                     hipMemcpyDeviceToDevice, st),
                      "solve: update");
            }
            check(oosg::launchScaleRows(R, sweepVectors, scale, n, k, st), "solve: scaling");
            sweeps(f, k, 0);
            if (scale != nullptr)
                check(oosg::launchScaleRows(sweepVectors, sweepVectors, scale, n, k, st), "solve: scaling");
            check(oosg::launchAddRows(X, sweepVectors, n, k, st), "solve: update");
        }
        const int slot = p.profiler.begin(PROFILE_RESIDUAL, residualBytes(m, n, k), st);
        check(oosg::launchMultiplyRows(m.dCsrOffsets.as<int64_t>(), m.dCsrCols.as<int32_t>(), m.dCsrValues.as<double>(),
                                       m.dCsrGroups.as<int32_t>(), static_cast<int64_t>(m.csrGroups.size()) - 1, X,
                                       B, R, k, st),
              "solve: residual");
        p.profiler.end(slot, st);
        startColumnMax(p, R, k, 1);
        if (step == 1)
        {
            // The first solution into x at once, its column maxima taken in the same pass (a later step writes x
            // again).
            unsigned long long* xNorm = p.solveNorms.as<unsigned long long>() + 32;
            check(oosg::launchFillInt32(reinterpret_cast<int32_t*>(xNorm), 32, 0, st), "solve: norms");
            check(oosg::launchFromRowMajor(X, x, ld, n, k, nullptr, 0, st, xNorm), "solve: layout");
            written = true;
        }
        else
        {
            startColumnMax(p, X, k, 2);
            written = false;
        }
        const std::vector<std::vector<double>> norms = finishColumnMax(p, 3, k);
        if (step == 1)
            bNorm = norms[0];
        const std::vector<double>& rNorm = norms[1];
        const std::vector<double>& xNorm = norms[2];
        double eta = 0.0;
        for (int h = 0; h < k; ++h)
        {
            const double denominator = m.normInf * xNorm[h] + bNorm[h];
            const double e = denominator > 0 ? rNorm[h] / denominator : rNorm[h];
            eta = std::max(eta, e != e ? INFINITY : e);
        }
        report.steps = step;
        report.backward_error = eta;
        if (eta <= target)
        {
            report.converged = 1;
            break;
        }
        if (step > 1 && !(eta < 0.5 * previous))
            break;
        previous = eta;
    }
    if (!written)
        check(oosg::launchFromRowMajor(X, x, ld, n, k, nullptr, 0, st), "solve: layout");
}

} // namespace

extern "C" {

const char* oosg_last_error(void)
{
    return g_lastError.c_str();
}

oosg_status oosg_last_error_equations(int32_t* equations, int64_t capacity, int64_t* count)
{
    const std::vector<int32_t> kept = g_lastEquations; // the calls below must not replace them
    if (count != nullptr)
        *count = static_cast<int64_t>(kept.size());
    if (equations != nullptr && capacity > 0)
        std::copy(kept.begin(), kept.begin() + std::min<int64_t>(capacity, static_cast<int64_t>(kept.size())),
                  equations);
    return OOSG_OK;
}

oosg_status oosg_last_error_shortfall(oosg_memory_shortfall* shortfall)
{
    if (shortfall == nullptr)
        return OOSG_INVALID_ARGUMENT;
    *shortfall = g_lastShortfall;
    return OOSG_OK;
}

oosg_status oosg_csc_checksum(const oosg_csc_desc* matrix, uint64_t* checksum)
{
    return guarded(
        [&]
        {
            if (matrix == nullptr || checksum == nullptr || matrix->n < 0 ||
                (matrix->n > 0 && (matrix->offsets == nullptr || matrix->rows == nullptr || matrix->values == nullptr)))
                fail(OOSG_INVALID_ARGUMENT, "checksum: null argument");
            // FNV-1a (Fowler, Noll, Vo), 64 bits, one word at a time.
            constexpr uint64_t OFFSET_BASIS = 14695981039346656037ULL, PRIME = 1099511628211ULL;
            uint64_t hash = OFFSET_BASIS;
            const auto mix = [&hash](uint64_t word)
            {
                hash ^= word;
                hash *= PRIME;
            };
            const int64_t n = matrix->n, nnz = n > 0 ? matrix->offsets[n] : 0;
            mix(static_cast<uint64_t>(n));
            for (int64_t j = 0; j <= n && n > 0; ++j)
                mix(static_cast<uint64_t>(matrix->offsets[j]));
            if (n == 0)
                mix(0); // the one offset of an empty matrix
            for (int64_t e = 0; e < nnz; ++e)
                mix(static_cast<uint32_t>(matrix->rows[e]));
            for (int64_t e = 0; e < nnz; ++e)
            {
                uint64_t bits = 0;
                std::memcpy(&bits, &matrix->values[e], sizeof(bits));
                mix(bits);
            }
            *checksum = hash;
        });
}

oosg_status oosg_estimate_memory_bytes(const oosg_symbolic_desc* symbolic, int32_t pivoting, uint64_t* bytes)
{
    return guarded(
        [&]
        {
            if (symbolic == nullptr || bytes == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "estimate: null argument");
            if (pivoting != OOSG_POSITIVE_DEFINITE && pivoting != OOSG_INDEFINITE)
                fail(OOSG_INVALID_ARGUMENT, "estimate: unknown pivoting mode");
            // The plan's host part only: the structure and the schedule need no device.
            oosg_plan p;
            buildStructure(p, *symbolic);
            buildSchedule(p);
            *bytes = planBytes(p);
        });
}

oosg_status oosg_plan_memory(const oosg_plan* plan, uint64_t* reserved_bytes, uint64_t* budget_bytes)
{
    return guarded(
        [&]
        {
            if (plan == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan: null argument");
            if (reserved_bytes != nullptr)
                *reserved_bytes = plan->reserved;
            if (budget_bytes != nullptr)
                *budget_bytes = plan->budget;
        });
}

oosg_status oosg_plan_set_stream(oosg_plan* plan, void* stream)
{
    return guarded(
        [&]
        {
            if (plan == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan: null argument");
            plan->callerStream = static_cast<hipStream_t>(stream);
        });
}

oosg_status oosg_plan_set_cancellation(oosg_plan* plan, oosg_cancel_fn cancelled, void* context)
{
    return guarded(
        [&]
        {
            if (plan == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan: null argument");
            plan->cancelled = cancelled;
            plan->cancelContext = context;
        });
}

oosg_status oosg_plan_create(const oosg_symbolic_desc* symbolic, int32_t device, uint64_t hbm_budget_bytes,
                             oosg_plan** out)
{
    return oosg_plan_create_partial(symbolic, 0, device, hbm_budget_bytes, out);
}

oosg_status oosg_plan_create_partial(const oosg_symbolic_desc* symbolic, int32_t trailing_dof_count, int32_t device,
                                     uint64_t hbm_budget_bytes, oosg_plan** out)
{
    return createPlan(symbolic, trailing_dof_count, false, device, hbm_budget_bytes, out);
}

oosg_status oosg_plan_create_out_of_core(const oosg_symbolic_desc* symbolic, int32_t device, uint64_t hbm_budget_bytes,
                                         oosg_plan** out)
{
    return createPlan(symbolic, 0, true, device, hbm_budget_bytes, out);
}

} // extern "C"

namespace
{

oosg_status createPlan(const oosg_symbolic_desc* symbolic, int32_t trailing_dof_count, bool outOfCore, int32_t device,
                       uint64_t hbm_budget_bytes, oosg_plan** out)
{
    return guarded(
        [&]
        {
            if (symbolic == nullptr || out == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan: null argument");
            *out = nullptr;
            if (trailing_dof_count < 0)
                fail(OOSG_INVALID_ARGUMENT, "plan: a negative trailing count");
            auto p = std::make_unique<oosg_plan>();
            p->device = device;
            p->trailingDofs = trailing_dof_count;
            p->outOfCore = outOfCore;
            int count = 0;
            check(hipGetDeviceCount(&count), "hipGetDeviceCount");
            if (device < 0 || device >= count)
                fail(OOSG_INVALID_ARGUMENT, "plan: no device " + std::to_string(device));
            setDevice(p.get());
            buildStructure(*p, *symbolic);
            buildSchedule(*p);
            buildTrailing(*p);
            p->reserved = structureBytes(*p); // uploaded uncounted below, held while the plan lives
            // The memory of a factorisation and of its solves, reserved before any numeric work (MemoryBudgetError).
            size_t freeBytes = 0, totalBytes = 0;
            check(hipMemGetInfo(&freeBytes, &totalBytes), "hipMemGetInfo");
            p->budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : freeBytes;
            const uint64_t need = planBytes(*p);
            if (need > p->budget)
            {
                const uint64_t panels = static_cast<uint64_t>(p->panelOff[p->S]) * 8;
                const uint64_t updates = static_cast<uint64_t>(p->updateOff[p->S]) * 8;
                Failure f{OOSG_OUT_OF_MEMORY,
                          "plan: the factorisation and its solves need " + mebibytes(need) +
                              " of device memory (panels " + mebibytes(panels) + ", update matrices " +
                              mebibytes(updates) + ", inverses, pivots and the solves' vectors " +
                              mebibytes(need - panels - updates) + "); the budget is " + mebibytes(p->budget),
                          {},
                          {}};
                f.shortfall.requested_bytes = need;
                f.shortfall.available_bytes = p->budget;
                f.shortfall.limit_bytes = p->budget;
                throw f;
            }
            check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
            p->lanes[LANE_CRITICAL] = p->stream;
            for (int lane = 1; lane < LANES; ++lane)
                check(hipStreamCreateWithFlags(&p->lanes[lane], hipStreamNonBlocking), "hipStreamCreate");
            for (hipEvent_t& e : p->laneEvents)
                check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
            check(oosg::configureKernels(), "plan: kernel limits");
            check(hipEventCreateWithFlags(&p->callerEvent, hipEventDisableTiming), "hipEventCreate");
            check(oosg::scratchAddress(device, &p->scratch), "plan: the scratch array's address");
            check(hipEventCreate(&p->start), "hipEventCreate");
            check(hipEventCreate(&p->stop), "hipEventCreate");
            if (p->outOfCore)
            {
                check(hipStreamCreateWithFlags(&p->copyStream, hipStreamNonBlocking), "hipStreamCreate");
                for (int w = 0; w < 2; ++w)
                {
                    check(hipEventCreateWithFlags(&p->oocReady[w], hipEventDisableTiming), "hipEventCreate");
                    check(hipEventCreateWithFlags(&p->oocFree[w], hipEventDisableTiming), "hipEventCreate");
                }
                // The windows' launches run one after the other on the plan's stream, launch by launch.
                p->graphs = false;
            }
            uploadPlan(*p);
            reserveSolveBuffers(*p);
            p->storagePool.push_back(takeStorage(*p));
            *out = p.release();
        });
}

} // namespace

extern "C" {

void oosg_plan_destroy(oosg_plan* plan)
{
    if (plan == nullptr)
        return;
    releasePlan(plan);
}

oosg_status oosg_plan_stats(const oosg_plan* plan, int64_t* stats)
{
    return guarded(
        [&]
        {
            if (plan == nullptr || stats == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan stats: null argument");
            stats[0] = plan->S;
            stats[1] = plan->levels;
            stats[2] = plan->panelOff[plan->S];
            stats[3] = plan->updateOff[plan->S];
            stats[4] = plan->smallFronts;
            stats[5] = plan->largeFronts;
            stats[6] = std::count_if(plan->schedule.begin(), plan->schedule.end(),
                                     [](const Launch& l) { return l.kind != LaunchKind::Sync; });
            stats[7] = static_cast<int64_t>(plan->tiles.size());
        });
}

oosg_status oosg_plan_profile(oosg_plan* plan, int32_t enable)
{
    return guarded(
        [&]
        {
            if (plan == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "profile: null plan");
            plan->profiler.enabled = enable != 0;
            plan->profiler.reset();
        });
}

oosg_status oosg_plan_profile_read(const oosg_plan* plan, double* seconds, int64_t* launches, double* work)
{
    return guarded(
        [&]
        {
            if (plan == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "profile: null plan");
            for (int k = 0; k < PROFILE_KINDS; ++k)
            {
                if (seconds)
                    seconds[k] = plan->profiler.seconds[k];
                if (launches)
                    launches[k] = plan->profiler.launches[k];
                if (work)
                    work[k] = plan->profiler.work[k];
            }
        });
}

// ---- the learned precision selector ------------------------------------------------------------------------------

} // extern "C"

struct oosg_selector
{
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceBuffer params, m, v, gradient, loss;
    int step = 0; // Adam steps taken
    ~oosg_selector()
    {
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

namespace
{

// Device work arrays of a batch: features, labels, activations, deltas.
struct SelectorBatch
{
    DeviceBuffer x, y, a1, a2, p, z, d3, d2, d1;
    SelectorBatch(int64_t count, hipStream_t stream, const float* features, const float* labels)
    {
        const size_t n = static_cast<size_t>(count);
        const size_t h = static_cast<size_t>(oosg::SELECTOR_HIDDEN);
        x.allocate(n * oosg::SELECTOR_FEATURES * sizeof(float), "selector: features");
        check(hipMemcpyAsync(x.ptr, features, n * oosg::SELECTOR_FEATURES * sizeof(float), hipMemcpyHostToDevice,
                             stream),
              "selector: features");
        if (labels != nullptr)
        {
            y.allocate(n * sizeof(float), "selector: labels");
            check(hipMemcpyAsync(y.ptr, labels, n * sizeof(float), hipMemcpyHostToDevice, stream), "selector: labels");
            a1.allocate(n * h * sizeof(float), "selector: activations");
            a2.allocate(n * h * sizeof(float), "selector: activations");
            d3.allocate(n * sizeof(float), "selector: deltas");
            d2.allocate(n * h * sizeof(float), "selector: deltas");
            d1.allocate(n * h * sizeof(float), "selector: deltas");
        }
        p.allocate(n * sizeof(float), "selector: probabilities");
        z.allocate(n * sizeof(float), "selector: logits");
    }
};

void checkBatch(const oosg_selector* s, const float* features, const float* labels, int64_t count, bool needLabels)
{
    if (s == nullptr || features == nullptr || count <= 0 || (needLabels && labels == nullptr))
        fail(OOSG_INVALID_ARGUMENT, "selector: null argument or empty batch");
    if (needLabels)
        for (int64_t b = 0; b < count; ++b)
            if (!(labels[b] >= 0.0f && labels[b] <= 1.0f))
                fail(OOSG_INVALID_ARGUMENT, "selector: labels lie in [0, 1]");
}

} // namespace

extern "C" {

oosg_status oosg_selector_create(int32_t device, uint64_t seed, oosg_selector** out)
{
    return guarded(
        [&]
        {
            if (out == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "selector: null argument");
            *out = nullptr;
            auto s = std::make_unique<oosg_selector>();
            s->device = device;
            check(hipSetDevice(device), "hipSetDevice");
            check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
            const size_t bytes = oosg::SELECTOR_PARAMETERS * sizeof(float);
            s->params.allocate(bytes, "selector: parameters");
            s->m.allocate(bytes, "selector: moments");
            s->v.allocate(bytes, "selector: moments");
            s->gradient.allocate(bytes, "selector: gradient");
            s->loss.allocate(sizeof(float), "selector: loss");
            // Glorot-uniform weights from a splitmix64 sequence, zero biases.
            std::vector<float> w(oosg::SELECTOR_PARAMETERS, 0.0f);
            uint64_t state = seed;
            const auto uniform = [&state]
            {
                uint64_t z = (state += 0x9E3779B97F4A7C15ull);
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                return static_cast<double>(z >> 11) * 0x1.0p-53;
            };
            const int F = oosg::SELECTOR_FEATURES, H = oosg::SELECTOR_HIDDEN;
            const auto fill = [&](int offset, int count, int fanIn, int fanOut)
            {
                const double a = std::sqrt(6.0 / (fanIn + fanOut));
                for (int q = 0; q < count; ++q)
                    w[offset + q] = static_cast<float>((2.0 * uniform() - 1.0) * a);
            };
            fill(0, H * F, F, H);
            fill(H * F + H, H * H, H, H);
            fill(H * F + H + H * H + H, H, H, 1);
            check(hipMemcpyAsync(s->params.ptr, w.data(), bytes, hipMemcpyHostToDevice, s->stream), "selector");
            std::vector<float> zero(oosg::SELECTOR_PARAMETERS, 0.0f);
            check(hipMemcpyAsync(s->m.ptr, zero.data(), bytes, hipMemcpyHostToDevice, s->stream), "selector");
            check(hipMemcpyAsync(s->v.ptr, zero.data(), bytes, hipMemcpyHostToDevice, s->stream), "selector");
            check(hipStreamSynchronize(s->stream), "selector");
            *out = s.release();
        });
}

void oosg_selector_destroy(oosg_selector* selector)
{
    if (selector == nullptr)
        return;
    (void)hipSetDevice(selector->device);
    delete selector;
}

oosg_status oosg_selector_parameters(const oosg_selector* s, float* params)
{
    return guarded(
        [&]
        {
            if (s == nullptr || params == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "selector: null argument");
            check(hipSetDevice(s->device), "hipSetDevice");
            check(hipMemcpyAsync(params, s->params.ptr, oosg::SELECTOR_PARAMETERS * sizeof(float),
                                 hipMemcpyDeviceToHost, s->stream),
                  "selector: parameters");
            check(hipStreamSynchronize(s->stream), "selector");
        });
}

oosg_status oosg_selector_set_parameters(oosg_selector* s, const float* params)
{
    return guarded(
        [&]
        {
            if (s == nullptr || params == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "selector: null argument");
            check(hipSetDevice(s->device), "hipSetDevice");
            check(hipMemcpyAsync(s->params.ptr, params, oosg::SELECTOR_PARAMETERS * sizeof(float),
                                 hipMemcpyHostToDevice, s->stream),
                  "selector: parameters");
            check(hipStreamSynchronize(s->stream), "selector");
            s->step = 0;
        });
}

// hostValues of a matrix whose values were last set on the device.
static void refreshHostValues(const oosg_matrix& cm)
{
    oosg_matrix& m = const_cast<oosg_matrix&>(cm);
    if (!m.hostStale)
        return;
    setDevice(m.plan);
    m.hostValues.resize(static_cast<size_t>(m.nnz));
    check(hipMemcpy(m.hostValues.data(), m.dValues.ptr, m.hostValues.size() * sizeof(double), hipMemcpyDeviceToHost),
          "matrix: values");
    m.hostStale = false;
}

oosg_status oosg_selector_cut_features(const oosg_plan* plan, const oosg_matrix* m, float* features,
                                       int32_t capacity, int32_t* count)
{
    return guarded(
        [&]
        {
            if (plan == nullptr || m == nullptr || count == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "selector: null argument");
            if (m->plan != plan)
                fail(OOSG_INVALID_ARGUMENT, "selector: the matrix was made for another plan");
            const int32_t L = plan->levels;
            *count = L;
            if (features == nullptr)
                return;
            if (capacity < L)
                fail(OOSG_INVALID_ARGUMENT, "selector: room for " + std::to_string(L) + " cuts (8 features each) needed");
            refreshHostValues(*m);
            setDevice(plan);
            const int32_t n = m->n;
            // The scaled matrix's off-diagonal row sums and its diagonal dominance per equation.
            std::vector<double> scale(static_cast<size_t>(n), 1.0);
            if (n > 0 && m->dScale.ptr != nullptr)
                check(hipMemcpy(scale.data(), m->dScale.ptr, scale.size() * sizeof(double), hipMemcpyDeviceToHost),
                      "selector: scales");
            std::vector<double> off(static_cast<size_t>(n), 0.0), diagonal(static_cast<size_t>(n), 0.0);
            for (int32_t j = 0; j < n; ++j)
                for (int64_t e = m->offsets[j]; e < m->offsets[j + 1]; ++e)
                {
                    const int32_t i = m->rows[e];
                    const double a = std::fabs(m->hostValues[e]) * scale[i] * scale[j];
                    if (i == j)
                    {
                        diagonal[j] = a;
                        continue;
                    }
                    off[i] += a;
                    off[j] += a;
                }
            std::vector<float> growth(static_cast<size_t>(plan->S), 0.0f);
            if (plan->growthRecorded && plan->S > 0)
                check(hipMemcpy(growth.data(), plan->supernodeGrowth.ptr, growth.size() * sizeof(float),
                                hipMemcpyDeviceToHost),
                      "selector: pivot growth");
            // Per level: flops, the largest front order, the dominance sum, minimum and count, the largest growth.
            std::vector<double> flops(L, 0.0), order(L, 0.0), domSum(L, 0.0), domMin(L, 3.0), growthMax(L, 0.0);
            std::vector<int64_t> equations(L, 0);
            double totalFlops = 0.0;
            for (int32_t s = 0; s < plan->S; ++s)
            {
                const int32_t l = plan->levelOf[s];
                if (l < 0 || l >= L)
                    continue;
                const double c = plan->ncol[s], r = plan->nrow[s];
                const double work = c * c * c / 3.0 + c * c * r + c * r * r;
                flops[l] += work;
                totalFlops += work;
                order[l] = std::max(order[l], c + r);
                growthMax[l] = std::max(growthMax[l], static_cast<double>(growth[s]));
                for (int32_t j = plan->first[s]; j < plan->first[s] + plan->ncol[s]; ++j)
                {
                    double d = off[j] > 0.0 ? std::log10(diagonal[j] / off[j]) : 3.0;
                    d = d != d ? -3.0 : std::min(3.0, std::max(-3.0, d));
                    domSum[l] += d;
                    domMin[l] = std::min(domMin[l], d);
                    ++equations[l];
                }
            }
            std::vector<double> orderAbove(static_cast<size_t>(L) + 1, 0.0);
            for (int32_t l = L - 1; l >= 0; --l)
                orderAbove[l] = std::max(orderAbove[l + 1], order[l]);
            double flopsBelow = 0.0, orderBelow = 1.0, sumBelow = 0.0, minBelow = 3.0, growthBelow = 0.0;
            int64_t equationsBelow = 0;
            for (int32_t cut = 1; cut <= L; ++cut)
            {
                const int32_t l = cut - 1; // the level the cut adds to the FP32 part
                flopsBelow += flops[l];
                orderBelow = std::max(orderBelow, order[l]);
                sumBelow += domSum[l];
                minBelow = std::min(minBelow, domMin[l]);
                equationsBelow += equations[l];
                growthBelow = std::max(growthBelow, growthMax[l]);
                float* f = features + 8 * static_cast<size_t>(l);
                f[0] = static_cast<float>(static_cast<double>(cut) / L);
                f[1] = static_cast<float>(totalFlops > 0.0 ? flopsBelow / totalFlops : 0.0);
                f[2] = static_cast<float>(std::log10(orderBelow) - 2.5);
                f[3] = static_cast<float>(std::log10(std::max(orderAbove[cut], 1.0)) - 2.5);
                f[4] = static_cast<float>(std::log10(std::max<double>(n, 1)) - 5.0);
                f[5] = static_cast<float>(equationsBelow > 0 ? sumBelow / static_cast<double>(equationsBelow) : 0.0);
                f[6] = static_cast<float>(minBelow);
                f[7] = static_cast<float>(std::isfinite(growthBelow) ? std::log10(1.0 + growthBelow) / 4.0 : 10.0);
            }
        });
}

oosg_status oosg_selector_predict(oosg_selector* s, const float* features, int64_t count, float* probability)
{
    return guarded(
        [&]
        {
            checkBatch(s, features, nullptr, count, false);
            if (probability == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "selector: null argument");
            check(hipSetDevice(s->device), "hipSetDevice");
            SelectorBatch batch(count, s->stream, features, nullptr);
            check(oosg::selectorForward(s->params.as<float>(), batch.x.as<float>(), count, nullptr, nullptr,
                                        batch.p.as<float>(), nullptr, s->stream),
                  "selector: forward");
            check(hipMemcpyAsync(probability, batch.p.ptr, static_cast<size_t>(count) * sizeof(float),
                                 hipMemcpyDeviceToHost, s->stream),
                  "selector: forward");
            check(hipStreamSynchronize(s->stream), "selector");
        });
}

oosg_status oosg_selector_gradient(oosg_selector* s, const float* features, const float* labels, int64_t count,
                                   float* loss, float* gradient)
{
    return guarded(
        [&]
        {
            checkBatch(s, features, labels, count, true);
            check(hipSetDevice(s->device), "hipSetDevice");
            SelectorBatch b(count, s->stream, features, labels);
            check(oosg::selectorGradient(s->params.as<float>(), b.x.as<float>(), b.y.as<float>(), count,
                                         b.a1.as<float>(), b.a2.as<float>(), b.p.as<float>(), b.z.as<float>(),
                                         b.d3.as<float>(), b.d2.as<float>(), b.d1.as<float>(),
                                         s->gradient.as<float>(), s->loss.as<float>(), s->stream),
                  "selector: gradient");
            if (gradient != nullptr)
                check(hipMemcpyAsync(gradient, s->gradient.ptr, oosg::SELECTOR_PARAMETERS * sizeof(float),
                                     hipMemcpyDeviceToHost, s->stream),
                      "selector: gradient");
            if (loss != nullptr)
                check(hipMemcpyAsync(loss, s->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, s->stream),
                      "selector: loss");
            check(hipStreamSynchronize(s->stream), "selector");
        });
}

oosg_status oosg_selector_train(oosg_selector* s, const float* features, const float* labels, int64_t count,
                                int32_t steps, float learning_rate, float* loss)
{
    return guarded(
        [&]
        {
            checkBatch(s, features, labels, count, true);
            if (steps < 1 || !(learning_rate > 0.0f))
                fail(OOSG_INVALID_ARGUMENT, "selector: steps >= 1 and a positive learning rate");
            check(hipSetDevice(s->device), "hipSetDevice");
            SelectorBatch b(count, s->stream, features, labels);
            for (int32_t q = 0; q < steps; ++q)
            {
                check(oosg::selectorGradient(s->params.as<float>(), b.x.as<float>(), b.y.as<float>(), count,
                                             b.a1.as<float>(), b.a2.as<float>(), b.p.as<float>(), b.z.as<float>(),
                                             b.d3.as<float>(), b.d2.as<float>(), b.d1.as<float>(),
                                             s->gradient.as<float>(), s->loss.as<float>(), s->stream),
                      "selector: gradient");
                check(oosg::selectorAdam(s->params.as<float>(), s->gradient.as<float>(), s->m.as<float>(),
                                         s->v.as<float>(), learning_rate, ++s->step, s->stream),
                      "selector: step");
            }
            if (loss != nullptr)
                check(hipMemcpyAsync(loss, s->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, s->stream),
                      "selector: loss");
            check(hipStreamSynchronize(s->stream), "selector");
        });
}

oosg_status oosg_plan_launch_times(const oosg_plan* plan, int32_t* count, int32_t* kinds, int32_t* levels, float* ms,
                                   double* flops)
{
    return guarded(
        [&]
        {
            if (plan == nullptr || count == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan: null argument");
            const int32_t n = static_cast<int32_t>(plan->schedule.size());
            const bool measured = plan->launchMs.size() == plan->schedule.size();
            if (kinds != nullptr || levels != nullptr || ms != nullptr || flops != nullptr)
            {
                if (*count < n)
                    fail(OOSG_INVALID_ARGUMENT, "plan: launch arrays too short");
                for (int32_t q = 0; q < n; ++q)
                {
                    const Launch& l = plan->schedule[q];
                    if (kinds != nullptr)
                        kinds[q] = l.profile;
                    if (levels != nullptr)
                        levels[q] = l.level;
                    if (ms != nullptr)
                        ms[q] = measured ? plan->launchMs[q] : -1.0f;
                    if (flops != nullptr)
                        flops[q] = l.flops;
                }
            }
            *count = n;
        });
}

oosg_status oosg_plan_streams(oosg_plan* plan, int32_t precision, int32_t count)
{
    return guarded(
        [&]
        {
            if (plan == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan: null argument");
            if (precision != OOSG_FP64 && precision != OOSG_FP32)
                fail(OOSG_INVALID_ARGUMENT, "plan: unknown precision");
            if (count != 1 && count != LANES)
                fail(OOSG_INVALID_ARGUMENT, "plan: the factorisation runs on 1 or 3 streams");
            setDevice(plan);
            plan->concurrent[precision] = count == LANES;
            plan->eager[precision] = false;
            plan->factorGraph[precision].clear();
        });
}

oosg_status oosg_plan_launch_mode(oosg_plan* plan, int32_t precision, int32_t mode)
{
    return guarded(
        [&]
        {
            if (plan == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan: null argument");
            if (precision != OOSG_FP64 && precision != OOSG_FP32)
                fail(OOSG_INVALID_ARGUMENT, "plan: unknown precision");
            if (mode < OOSG_LAUNCH_GRAPH || mode > OOSG_LAUNCH_EAGER_STREAMS)
                fail(OOSG_INVALID_ARGUMENT, "plan: unknown launch mode");
            setDevice(plan);
            plan->concurrent[precision] = mode != OOSG_LAUNCH_GRAPH;
            plan->eager[precision] = mode == OOSG_LAUNCH_EAGER_STREAMS;
            plan->factorGraph[precision].clear();
        });
}

oosg_status oosg_plan_get_launch_mode(const oosg_plan* plan, int32_t precision, int32_t* mode)
{
    return guarded(
        [&]
        {
            if (plan == nullptr || mode == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan: null argument");
            if (precision != OOSG_FP64 && precision != OOSG_FP32)
                fail(OOSG_INVALID_ARGUMENT, "plan: unknown precision");
            *mode = !plan->concurrent[precision] ? OOSG_LAUNCH_GRAPH
                    : plan->eager[precision]     ? OOSG_LAUNCH_EAGER_STREAMS
                                                 : OOSG_LAUNCH_GRAPH_STREAMS;
        });
}

oosg_status oosg_plan_tile_policy(oosg_plan* plan, int32_t policy)
{
    return guarded(
        [&]
        {
            if (plan == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan: null argument");
            if (policy < 0 || policy > 2)
                fail(OOSG_INVALID_ARGUMENT, "plan: the tile policy is 0 (64 x 64), 1 (128 x 128) or 2 (by level)");
            setDevice(plan);
            if (plan->tilePolicy != policy)
            {
                plan->tilePolicy = policy;
                plan->factorGraph[0].clear(); // the captured launches change
                plan->factorGraph[1].clear();
            }
        });
}

oosg_status oosg_plan_autotune(oosg_plan* plan, const oosg_matrix* matrix, int32_t precision, int32_t* big_launches,
                               int32_t* tuned_launches, int32_t* streams)
{
    return guarded(
        [&]
        {
            if (plan == nullptr || matrix == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "autotune: null argument");
            if (matrix->plan != plan)
                fail(OOSG_INVALID_ARGUMENT, "autotune: the matrix was made for another plan");
            if (precision != OOSG_FP64 && precision != OOSG_FP32)
                fail(OOSG_INVALID_ARGUMENT, "autotune: unknown precision");
            if (plan->outOfCore)
                fail(OOSG_INVALID_ARGUMENT, "autotune: an out-of-core plan launches one level at a time on one "
                                            "stream; there is nothing to tune");
            setDevice(plan);
            oosg_plan& p = *plan;
            // A factor's storage of the plan's (given back at the end), and the plan's pivot status.
            struct Lent
            {
                oosg_plan& p;
                FactorStorage storage;
                ~Lent()
                {
                    if (p.storagePool.size() < 2)
                        p.storagePool.push_back(std::move(storage));
                }
            } lent{p, takeStorage(p)};
            DeviceBuffer& panels = lent.storage.panels;
            DeviceBuffer& linv = lent.storage.linv;
            DeviceBuffer& diag = lent.storage.diag;
            DeviceBuffer& zero = p.pivotStatus;
            const int32_t keep = p.tilePolicy;
            const bool profiling = p.profiler.enabled;
            p.profiler.enabled = false;
            // Each policy's per-launch times: a warm run, then the measured one (eager launches, no graph).
            const auto measure = [&](int32_t policy)
            {
                p.tilePolicy = policy;
                p.timeLaunches = true;
                std::vector<float> best;
                for (int run = 0; run < 2; ++run)
                {
                    check(oosg::launchFillInt32(zero.as<int32_t>(), p.S, INT32_MAX, p.stream), "autotune: status");
                    if (precision == OOSG_FP32)
                        runFactor<float>(p, *matrix, panels.as<float>(), linv.as<float>(), diag.as<double>(),
                                         zero.as<int32_t>(), true);
                    else
                        runFactor<double>(p, *matrix, panels.as<double>(), linv.as<double>(), diag.as<double>(),
                                          zero.as<int32_t>(), false);
                    best = p.launchMs;
                }
                p.timeLaunches = false;
                return best;
            };
            std::vector<float> small, big;
            try
            {
                small = measure(0);
                big = measure(1);
            }
            catch (...)
            {
                p.tilePolicy = keep;
                p.timeLaunches = false;
                p.profiler.enabled = profiling;
                throw;
            }
            p.profiler.enabled = profiling;
            int32_t chosen = 0, tuned = 0;
            for (size_t li = 0; li < p.schedule.size(); ++li)
            {
                Launch& l = p.schedule[li];
                if (l.bigCount == 0)
                    continue;
                ++tuned;
                l.big[precision] = big[li] < small[li];
                chosen += l.big[precision] ? 1 : 0;
            }
            // The launch times as the kept policy runs them (the update matrices with the size chosen).
            p.launchMs = small;
            for (size_t li = 0; li < p.schedule.size() && li < big.size(); ++li)
                if (p.schedule[li].bigCount > 0 && p.schedule[li].big[precision])
                    p.launchMs[li] = big[li];
            p.tilePolicy = 2;
            p.tuned[precision] = true;
            // Then the whole factorisation in each launch mode, launched as a factorisation launches it (the best of
            // three runs each, after a warm one that captures the graph): one graph on the plan's stream, one graph
            // of three streams, three streams launched one launch at a time (where the runtime does not overlap a
            // graph's branches as it overlaps the streams). The fastest stays.
            const auto wall = [&](int32_t mode)
            {
                p.concurrent[precision] = mode != OOSG_LAUNCH_GRAPH;
                GraphCache trial;
                const auto body = [&]
                {
                    check(oosg::launchFillInt32(zero.as<int32_t>(), p.S, INT32_MAX, p.stream), "autotune: status");
                    if (precision == OOSG_FP32)
                        runFactor<float>(p, *matrix, panels.as<float>(), linv.as<float>(), diag.as<double>(),
                                         zero.as<int32_t>(), true);
                    else
                        runFactor<double>(p, *matrix, panels.as<double>(), linv.as<double>(), diag.as<double>(),
                                          zero.as<int32_t>(), false);
                };
                float best = INFINITY;
                for (int run = 0; run < 4; ++run)
                {
                    check(hipEventRecord(p.start, p.stream), "hipEventRecord");
                    if (p.graphs && mode != OOSG_LAUNCH_EAGER_STREAMS)
                        trial.run(p.stream, {}, body);
                    else
                        body();
                    check(hipEventRecord(p.stop, p.stream), "hipEventRecord");
                    check(hipEventSynchronize(p.stop), "autotune");
                    float ms = 0.0f;
                    check(hipEventElapsedTime(&ms, p.start, p.stop), "hipEventElapsedTime");
                    if (run > 0)
                        best = std::min(best, ms);
                }
                return best;
            };
            p.profiler.enabled = false;
            int32_t fastest = OOSG_LAUNCH_GRAPH;
            float fastestMs = wall(OOSG_LAUNCH_GRAPH);
            for (int32_t mode = OOSG_LAUNCH_GRAPH_STREAMS; mode <= (p.graphs ? OOSG_LAUNCH_EAGER_STREAMS : mode); ++mode)
            {
                const float ms = wall(mode);
                if (ms < fastestMs)
                {
                    fastest = mode;
                    fastestMs = ms;
                }
            }
            p.profiler.enabled = profiling;
            p.concurrent[precision] = fastest != OOSG_LAUNCH_GRAPH;
            p.eager[precision] = fastest == OOSG_LAUNCH_EAGER_STREAMS;
            p.factorGraph[0].clear();
            p.factorGraph[1].clear();
            if (big_launches != nullptr)
                *big_launches = chosen;
            if (tuned_launches != nullptr)
                *tuned_launches = tuned;
            if (streams != nullptr)
                *streams = p.concurrent[precision] ? LANES : 1;
        });
}

oosg_status oosg_plan_use_graphs(oosg_plan* plan, int32_t enable)
{
    return guarded(
        [&]
        {
            if (plan == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "plan: null argument");
            setDevice(plan);
            plan->graphs = enable != 0;
            if (!plan->graphs)
            {
                plan->factorGraph[0].clear();
                plan->factorGraph[1].clear();
                plan->sweepGraph.clear();
            }
        });
}

oosg_status oosg_plan_synchronize(oosg_plan* plan)
{
    return guarded(
        [&]
        {
            setDevice(plan);
            check(hipStreamSynchronize(plan->stream), "synchronize");
        });
}

oosg_status oosg_csc_from_blocks_device(int32_t device, const oosg_block_matrix_desc* blocks,
                                        const int32_t* equation_of_dof, int64_t dof_count, int32_t kept,
                                        int64_t* offsets, int32_t* rows, double* values, int64_t* nnz, void* stream)
{
    return guarded(
        [&]
        {
            if (blocks == nullptr || nnz == nullptr || offsets == nullptr || kept < 0 || blocks->node_count < 0 ||
                (blocks->node_count > 0 && (blocks->first_dof == nullptr || blocks->column_start == nullptr ||
                                            blocks->block_value_offsets == nullptr)) ||
                (dof_count > 0 && equation_of_dof == nullptr))
                fail(OOSG_INVALID_ARGUMENT, "SymmetricCsc::fromBlocks: null argument");
            int count = 0;
            check(hipGetDeviceCount(&count), "hipGetDeviceCount");
            if (device < 0 || device >= count)
                fail(OOSG_INVALID_ARGUMENT, "SymmetricCsc::fromBlocks: no device " + std::to_string(device));
            check(hipSetDevice(device), "hipSetDevice");
            const hipStream_t st = static_cast<hipStream_t>(stream);
            const int32_t N = blocks->node_count;
            // The sizes the arrays end with: the DOFs, the blocks, the values.
            int64_t totalDofs = 0, blockCount = 0, valueCount = 0;
            if (N > 0)
            {
                check(hipMemcpyAsync(&totalDofs, blocks->first_dof + N, sizeof(int64_t), hipMemcpyDeviceToHost, st),
                      "fromBlocks: sizes");
                check(hipMemcpyAsync(&blockCount, blocks->column_start + N, sizeof(int64_t), hipMemcpyDeviceToHost,
                                     st),
                      "fromBlocks: sizes");
                check(hipStreamSynchronize(st), "fromBlocks: sizes");
                if (blockCount < 0)
                    fail(OOSG_INVALID_ARGUMENT, "SymmetricCsc::fromBlocks: a negative block count");
                if (blockCount > 0 && blocks->block_rows == nullptr)
                    fail(OOSG_INVALID_ARGUMENT, "SymmetricCsc::fromBlocks: null argument");
                check(hipMemcpyAsync(&valueCount, blocks->block_value_offsets + blockCount, sizeof(int64_t),
                                     hipMemcpyDeviceToHost, st),
                      "fromBlocks: sizes");
                check(hipStreamSynchronize(st), "fromBlocks: sizes");
            }
            if (totalDofs != dof_count)
                // requireMatchingSelection (SymmetricCsc.cpp:25-37)
                fail(OOSG_INVALID_ARGUMENT,
                     "SymmetricCsc::fromBlocks: the DOF selection (" + std::to_string(dof_count) +
                         " DOFs) was not made for the pattern of the matrix (" + std::to_string(totalDofs) +
                         " DOFs); build it from the same node DOF counts");
            if (valueCount > 0 && blocks->values == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "SymmetricCsc::fromBlocks: null argument");
            const oosg::BlockMatrixView view{N, blocks->first_dof, blocks->column_start, blocks->block_rows,
                                             blocks->block_value_offsets, blocks->values};
            const auto scratch = [](DeviceBuffer& b, size_t bytes, const char* what)
            { b.allocate(std::max<size_t>(bytes, 8), what); };
            DeviceBuffer flag, hits, counts, pieceSums;
            scratch(flag, sizeof(int32_t), "fromBlocks: flag");
            scratch(hits, static_cast<size_t>(kept) * sizeof(int32_t), "fromBlocks: equations");
            scratch(counts, static_cast<size_t>(kept) * sizeof(int32_t), "fromBlocks: counts");
            scratch(pieceSums, static_cast<size_t>(oosg::scanPieces(kept)) * sizeof(int64_t), "fromBlocks: scan");
            check(oosg::launchFillInt32(flag.as<int32_t>(), 1, 0, st), "fromBlocks");
            check(oosg::launchFillInt32(hits.as<int32_t>(), kept, 0, st), "fromBlocks");
            check(oosg::launchFillInt32(counts.as<int32_t>(), kept, 0, st), "fromBlocks");
            check(oosg::launchBlocksValidate(view, equation_of_dof, dof_count, kept, blockCount, valueCount,
                                             hits.as<int32_t>(), flag.as<int32_t>(), st),
                  "fromBlocks: validation");
            int32_t bad = 0;
            check(hipMemcpyAsync(&bad, flag.ptr, sizeof(bad), hipMemcpyDeviceToHost, st), "fromBlocks: validation");
            check(hipStreamSynchronize(st), "fromBlocks: validation");
            if (bad != 0)
                fail(OOSG_INVALID_ARGUMENT,
                     std::string("SymmetricCsc::fromBlocks: the block matrix or the selection is not well formed:") +
                         ((bad & 1) ? " a node's DOF count is outside 1 to 6;" : "") +
                         ((bad & 2) ? " the block columns' starts do not ascend within the blocks;" : "") +
                         ((bad & 4) ? " a block column does not start with its diagonal block and go down by "
                                      "ascending rows;"
                                    : "") +
                         ((bad & 8) ? " a block's values do not span dofs(row) × dofs(column) within the values;"
                                    : "") +
                         ((bad & 16) ? " the equations are not a numbering of the kept DOFs;" : ""));
            check(oosg::launchBlockTerms(view, equation_of_dof, counts.as<int32_t>(), nullptr, nullptr, nullptr, st),
                  "fromBlocks: counts");
            check(oosg::launchExclusiveScan(counts.as<int32_t>(), kept, offsets, pieceSums.as<int64_t>(), st),
                  "fromBlocks: offsets");
            check(hipMemcpyAsync(nnz, offsets + kept, sizeof(int64_t), hipMemcpyDeviceToHost, st), "fromBlocks");
            check(hipStreamSynchronize(st), "fromBlocks: offsets");
            if (rows == nullptr)
                return;
            if (*nnz > 0 && values == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "SymmetricCsc::fromBlocks: null argument");
            DeviceBuffer cursor, scatteredRows, scatteredValues;
            scratch(cursor, static_cast<size_t>(kept) * sizeof(int64_t), "fromBlocks: slots");
            scratch(scatteredRows, static_cast<size_t>(*nnz) * sizeof(int32_t), "fromBlocks: terms");
            scratch(scatteredValues, static_cast<size_t>(*nnz) * sizeof(double), "fromBlocks: terms");
            if (kept > 0)
                check(hipMemcpyAsync(cursor.ptr, offsets, static_cast<size_t>(kept) * sizeof(int64_t),
                                     hipMemcpyDeviceToDevice, st),
                      "fromBlocks: slots");
            check(oosg::launchBlockTerms(view, equation_of_dof, nullptr, cursor.as<unsigned long long>(),
                                         scatteredRows.as<int32_t>(), scatteredValues.as<double>(), st),
                  "fromBlocks: terms");
            check(oosg::launchSortColumns(offsets, kept, scatteredRows.as<int32_t>(), scatteredValues.as<double>(),
                                          rows, values, st),
                  "fromBlocks: row order");
            check(hipStreamSynchronize(st), "fromBlocks");
        });
}

oosg_status oosg_matrix_create_device(oosg_plan* plan, int32_t n, const int64_t* offsets, const int32_t* rows,
                                      const double* values, oosg_matrix** out)
{
    std::vector<int64_t> hostOffsets;
    std::vector<int32_t> hostRows;
    std::vector<double> hostValues;
    const oosg_status copied = guarded(
        [&]
        {
            if (plan == nullptr || offsets == nullptr || out == nullptr || n < 0)
                fail(OOSG_INVALID_ARGUMENT, "matrix: null argument");
            setDevice(plan);
            hostOffsets.resize(static_cast<size_t>(n) + 1);
            check(hipMemcpy(hostOffsets.data(), offsets, hostOffsets.size() * sizeof(int64_t), hipMemcpyDeviceToHost),
                  "matrix: structure");
            const int64_t nnz = hostOffsets[n];
            if (hostOffsets[0] != 0 || nnz < 0)
                fail(OOSG_INVALID_ARGUMENT, "matrix: offsets[0] must be 0");
            if (nnz > 0 && (rows == nullptr || values == nullptr))
                fail(OOSG_INVALID_ARGUMENT, "matrix: null argument");
            hostRows.resize(static_cast<size_t>(nnz));
            hostValues.resize(static_cast<size_t>(nnz));
            if (nnz > 0)
            {
                check(hipMemcpy(hostRows.data(), rows, hostRows.size() * sizeof(int32_t), hipMemcpyDeviceToHost),
                      "matrix: structure");
                check(hipMemcpy(hostValues.data(), values, hostValues.size() * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "matrix: values");
            }
        });
    if (copied != OOSG_OK)
        return copied;
    const oosg_csc_desc desc{n, hostOffsets.data(), hostRows.data(), hostValues.data()};
    return oosg_matrix_create(plan, &desc, out);
}

oosg_status oosg_matrix_create(oosg_plan* plan, const oosg_csc_desc* csc, oosg_matrix** out)
{
    return guarded(
        [&]
        {
            if (plan == nullptr || csc == nullptr || out == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "matrix: null argument");
            *out = nullptr;
            if (csc->n != plan->n)
                fail(OOSG_INVALID_ARGUMENT, "supernodal factorisation: the matrix has " + std::to_string(csc->n) +
                                                " equations and the symbolic factor " + std::to_string(plan->n) +
                                                "; factorise the matrix assembled in the order of its symbolic analysis");
            setDevice(plan);
            auto m = std::make_unique<oosg_matrix>();
            m->plan = plan;
            const int32_t n = csc->n;
            m->n = n;
            m->offsets.assign(csc->offsets, csc->offsets + n + 1);
            m->nnz = m->offsets[n];
            if (m->offsets[0] != 0)
                fail(OOSG_INVALID_ARGUMENT, "matrix: offsets[0] must be 0");
            m->rows.assign(csc->rows, csc->rows + m->nnz);
            std::vector<double> values(csc->values, csc->values + m->nnz);
            // Front positions; the structure test of FrontAssembly.cpp:22-28.
            std::vector<int32_t> position(static_cast<size_t>(m->nnz)), column(static_cast<size_t>(m->nnz));
            m->diagonal.assign(n, 0.0);
            for (int32_t s = 0; s < plan->S; ++s)
            {
                const int32_t f = plan->first[s], c = plan->ncol[s];
                const int32_t* rws = plan->rows.data() + plan->rowStart[s];
                const int32_t r = plan->nrow[s];
                for (int32_t j = f; j < f + c; ++j)
                {
                    if (m->offsets[j + 1] < m->offsets[j])
                        fail(OOSG_INVALID_ARGUMENT, "matrix: column offsets must not decrease");
                    int32_t previous = -1;
                    for (int64_t k = m->offsets[j]; k < m->offsets[j + 1]; ++k)
                    {
                        const int32_t row = m->rows[k];
                        if (row <= previous)
                            fail(OOSG_INVALID_ARGUMENT, "matrix: the rows of a column must ascend");
                        previous = row;
                        int32_t pos = -1;
                        if (row >= j && row < f + c)
                            pos = row - f;
                        else if (row >= f + c)
                        {
                            const int32_t* it = std::lower_bound(rws, rws + r, row);
                            if (it != rws + r && *it == row)
                                pos = c + static_cast<int32_t>(it - rws);
                        }
                        if (pos < 0)
                            fail(OOSG_INVALID_ARGUMENT,
                                 "supernodal factorisation: the matrix entry (" + std::to_string(row) + ", " +
                                     std::to_string(j) + ") lies outside the structure of the symbolic factor");
                        position[k] = pos;
                        column[k] = j - f;
                        if (row == j)
                            m->diagonal[j] = values[k];
                    }
                }
            }
            // Power-of-two scales for the FP32 factor: s_j² a_jj in [1/2, 2) (no rounding: exact scaling).
            m->scale.assign(n, 1.0);
            for (int32_t j = 0; j < n; ++j)
            {
                const double a = m->diagonal[j];
                if (a > 0.0 && std::isfinite(a))
                {
                    int e = 0;
                    (void)std::frexp(a, &e);
                    m->scale[j] = std::ldexp(1.0, -(e / 2));
                }
            }
            // The full symmetric matrix by rows for the residuals.
            std::vector<int64_t> rowCount(static_cast<size_t>(n) + 1, 0);
            for (int32_t j = 0; j < n; ++j)
                for (int64_t k = m->offsets[j]; k < m->offsets[j + 1]; ++k)
                {
                    ++rowCount[m->rows[k] + 1];
                    if (m->rows[k] != j)
                        ++rowCount[j + 1];
                }
            for (int32_t i = 0; i < n; ++i)
                rowCount[i + 1] += rowCount[i];
            std::vector<int32_t> csrCols(static_cast<size_t>(rowCount[n]));
            m->csrFromCsc.resize(static_cast<size_t>(rowCount[n]));
            {
                std::vector<int64_t> fill(rowCount.begin(), rowCount.end() - 1);
                // Row i: the columns j < i (entries (i, j) of column j), then its own column from the diagonal.
                for (int32_t j = 0; j < n; ++j)
                    for (int64_t k = m->offsets[j]; k < m->offsets[j + 1]; ++k)
                    {
                        const int32_t i = m->rows[k];
                        if (i != j)
                        {
                            csrCols[fill[i]] = j;
                            m->csrFromCsc[fill[i]++] = k;
                        }
                    }
                for (int32_t j = 0; j < n; ++j)
                    for (int64_t k = m->offsets[j]; k < m->offsets[j + 1]; ++k)
                    {
                        csrCols[fill[j]] = m->rows[k];
                        m->csrFromCsc[fill[j]++] = k;
                    }
            }
            std::vector<double> csrValues(m->csrFromCsc.size());
            for (size_t e = 0; e < csrValues.size(); ++e)
                csrValues[e] = values[m->csrFromCsc[e]];
            m->normInf = 0.0;
            for (int32_t i = 0; i < n; ++i)
            {
                double sum = 0.0;
                for (int64_t e = rowCount[i]; e < rowCount[i + 1]; ++e)
                    sum += std::fabs(csrValues[e]);
                m->normInf = std::max(m->normInf, sum);
            }
            const hipStream_t st = plan->stream;
            // The matrix's device arrays are counted in the plan's budget (MemoryBudgetError before the upload).
            const uint64_t matrixBytes = m->offsets.size() * 8 + m->rows.size() * 4 + values.size() * 8 +
                                         position.size() * 4 + column.size() * 4 + m->scale.size() * 8 +
                                         m->diagonal.size() * 8 + rowCount.size() * 8 + csrCols.size() * 4 +
                                         csrValues.size() * 8 + (static_cast<uint64_t>(n) + 1) * 4;
            if (plan->reserved + matrixBytes > plan->budget)
                failBudget(*plan, matrixBytes, "the matrix (its lower and full forms, front positions)");
            m->dOffsets = upload(m->offsets, st, "matrix: offsets");
            m->dRows = upload(m->rows, st, "matrix: rows");
            m->dValues = upload(values, st, "matrix: values");
            m->hostValues = values;
            m->dPosition = upload(position, st, "matrix: positions");
            m->dColumn = upload(column, st, "matrix: columns");
            m->dScale = upload(m->scale, st, "matrix: scales");
            m->dDiagonal = upload(m->diagonal, st, "matrix: diagonal");
            m->dCsrOffsets = upload(rowCount, st, "matrix: row offsets");
            m->dCsrCols = upload(csrCols, st, "matrix: row columns");
            m->dCsrValues = upload(csrValues, st, "matrix: row values");
            // Row groups for the residual products.
            m->csrGroups.assign(1, 0);
            for (int32_t i = 1, first = 0; i <= n; ++i)
            {
                const bool same = i < n && i - first < 8 &&
                                  rowCount[i + 1] - rowCount[i] == rowCount[first + 1] - rowCount[first] &&
                                  std::equal(csrCols.begin() + rowCount[i], csrCols.begin() + rowCount[i + 1],
                                             csrCols.begin() + rowCount[first]);
                if (!same)
                {
                    m->csrGroups.push_back(i);
                    first = i;
                }
            }
            m->dCsrGroups = upload(m->csrGroups, st, "matrix: row groups");
            check(hipStreamSynchronize(st), "matrix: upload");
            for (DeviceBuffer* b : {&m->dOffsets, &m->dRows, &m->dValues, &m->dPosition, &m->dColumn, &m->dScale,
                                    &m->dDiagonal, &m->dCsrOffsets, &m->dCsrCols, &m->dCsrValues, &m->dCsrGroups})
            {
                b->account = &plan->reserved;
                plan->reserved += b->bytes;
            }
            holdPlan(plan);
            *out = m.release();
        });
}

oosg_status oosg_matrix_set_values(oosg_matrix* m, const double* values)
{
    return guarded(
        [&]
        {
            if (m == nullptr || values == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "matrix: null argument");
            setDevice(m->plan);
            const hipStream_t st = m->plan->stream;
            std::vector<double> v(values, values + m->nnz);
            for (int32_t j = 0; j < m->n; ++j)
            {
                m->diagonal[j] = 0.0;
                if (m->offsets[j] < m->offsets[j + 1] && m->rows[m->offsets[j]] == j)
                    m->diagonal[j] = v[m->offsets[j]];
                const double a = m->diagonal[j];
                m->scale[j] = 1.0;
                if (a > 0.0 && std::isfinite(a))
                {
                    int e = 0;
                    (void)std::frexp(a, &e);
                    m->scale[j] = std::ldexp(1.0, -(e / 2));
                }
            }
            std::vector<double> csrValues(m->csrFromCsc.size());
            for (size_t e = 0; e < csrValues.size(); ++e)
                csrValues[e] = v[m->csrFromCsc[e]];
            m->normInf = 0.0;
            std::vector<double> rowSum(m->n, 0.0);
            for (int32_t j = 0; j < m->n; ++j)
                for (int64_t k = m->offsets[j]; k < m->offsets[j + 1]; ++k)
                {
                    rowSum[m->rows[k]] += std::fabs(v[k]);
                    if (m->rows[k] != j)
                        rowSum[j] += std::fabs(v[k]);
                }
            for (const double s : rowSum)
                m->normInf = std::max(m->normInf, s);
            check(hipMemcpyAsync(m->dValues.ptr, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, st),
                  "matrix: values");
            check(hipMemcpyAsync(m->dScale.ptr, m->scale.data(), m->scale.size() * sizeof(double),
                                 hipMemcpyHostToDevice, st),
                  "matrix: scales");
            check(hipMemcpyAsync(m->dDiagonal.ptr, m->diagonal.data(), m->diagonal.size() * sizeof(double),
                                 hipMemcpyHostToDevice, st),
                  "matrix: diagonal");
            check(hipMemcpyAsync(m->dCsrValues.ptr, csrValues.data(), csrValues.size() * sizeof(double),
                                 hipMemcpyHostToDevice, st),
                  "matrix: row values");
            check(hipStreamSynchronize(st), "matrix: values");
            m->hostValues = std::move(v);
            m->hostStale = false;
        });
}

} // extern "C"

namespace
{

// The row map and the norm's slot of a matrix whose values change on the device (counted in the plan's budget).
void prepareDeviceValues(oosg_matrix& m)
{
    oosg_plan& p = *m.plan;
    if (m.dCsrMap.ptr != nullptr)
        return;
    const uint64_t bytes = m.csrFromCsc.size() * sizeof(int64_t) + sizeof(unsigned long long);
    if (p.reserved + bytes > p.budget)
        failBudget(p, bytes, "the matrix's row map (values set on the device)");
    m.dCsrMap = upload(m.csrFromCsc, p.stream, "matrix: row map");
    m.dCsrMap.account = &p.reserved;
    p.reserved += m.dCsrMap.bytes;
    allocateCounted(p, m.dNormKey, sizeof(unsigned long long), "matrix: norm");
}

// After new values in m.dValues (and its diagonal and scales): the rows' copy, the norm and the diagonal on the host;
// the host values are stale from here.
void valuesChangedOnDevice(oosg_matrix& m)
{
    oosg_plan& p = *m.plan;
    const hipStream_t st = p.stream;
    check(oosg::launchRowValues(m.n, m.dCsrOffsets.as<int64_t>(), m.dCsrMap.as<int64_t>(), m.dValues.as<double>(),
                                m.dCsrValues.as<double>(), m.dNormKey.as<unsigned long long>(), st),
          "matrix: rows");
    unsigned long long key = 0;
    check(hipMemcpyAsync(&key, m.dNormKey.ptr, sizeof(key), hipMemcpyDeviceToHost, st), "matrix: norm");
    check(hipMemcpyAsync(m.diagonal.data(), m.dDiagonal.ptr, m.diagonal.size() * sizeof(double),
                         hipMemcpyDeviceToHost, st),
          "matrix: diagonal");
    check(hipStreamSynchronize(st), "matrix: values");
    std::memcpy(&m.normInf, &key, sizeof(double));
    m.hostStale = true;
}

} // namespace

extern "C" {

oosg_status oosg_matrix_set_combination(oosg_matrix* m, double alpha, const oosg_matrix* a, double beta,
                                        const oosg_matrix* b, const double* scale)
{
    return guarded(
        [&]
        {
            if (m == nullptr || a == nullptr || b == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "matrix: null argument");
            if (a->plan != m->plan || b->plan != m->plan)
                fail(OOSG_INVALID_ARGUMENT, "SymmetricCsc::combine: the matrices were made for other plans");
            // The same structure: the plan fixes the positions, but not which of them hold entries.
            if (a->offsets != m->offsets || b->offsets != m->offsets || a->rows != m->rows || b->rows != m->rows)
                fail(OOSG_INVALID_ARGUMENT,
                     "SymmetricCsc::combine: the matrices must have the same structure, found " +
                         std::to_string(a->nnz) + " and " + std::to_string(b->nnz) +
                         " terms or other rows; assemble both on one sparse pattern and convert them with one DOF "
                         "selection");
            oosg_plan& p = *m->plan;
            setDevice(&p);
            followCaller(p);
            const hipStream_t st = p.stream;
            prepareDeviceValues(*m);
            DeviceBuffer dScaleIn;
            if (scale != nullptr)
            {
                std::vector<double> host(scale, scale + m->n);
                dScaleIn = upload(host, st, "matrix: scales");
                check(hipStreamSynchronize(st), "matrix: scales");
            }
            check(oosg::launchCombineValues(m->n, m->dOffsets.as<int64_t>(), m->dRows.as<int32_t>(), alpha,
                                            a->dValues.as<double>(), beta, b->dValues.as<double>(),
                                            dScaleIn.ptr != nullptr ? dScaleIn.as<double>() : nullptr,
                                            m->dValues.as<double>(), m->dDiagonal.as<double>(),
                                            m->dScale.as<double>(), st),
                  "matrix: combination");
            valuesChangedOnDevice(*m);
        });
}

oosg_status oosg_matrix_set_values_device(oosg_matrix* m, const double* values)
{
    return guarded(
        [&]
        {
            if (m == nullptr || (values == nullptr && m->nnz > 0))
                fail(OOSG_INVALID_ARGUMENT, "matrix: null argument");
            oosg_plan& p = *m->plan;
            setDevice(&p);
            followCaller(p);
            prepareDeviceValues(*m);
            check(oosg::launchCombineValues(m->n, m->dOffsets.as<int64_t>(), m->dRows.as<int32_t>(), 1.0, values, 0.0,
                                            nullptr, nullptr, m->dValues.as<double>(), m->dDiagonal.as<double>(),
                                            m->dScale.as<double>(), p.stream),
                  "matrix: values");
            valuesChangedOnDevice(*m);
        });
}

void oosg_matrix_destroy(oosg_matrix* m)
{
    if (m == nullptr)
        return;
    oosg_plan* plan = m->plan;
    (void)hipSetDevice(plan->device);
    delete m;
    releasePlan(plan);
    (void)hipGetLastError();
}

oosg_status oosg_factorise(oosg_plan* plan, const oosg_matrix* matrix, const oosg_factor_options* options,
                           oosg_factor** out)
{
    return guarded([&] { factoriseInto(plan, matrix, options, out); });
}

} // extern "C"

namespace
{

// A mixed factor's cut: the update matrices of the supernodes below it whose parents are above it, written in FP32
// and read in FP64, each widened in place through a staging buffer (the plan's, grown as needed): their ranges
// uploaded to the plan's widenRanges (whose bytes give their count).
void prepareWidening(oosg_plan& p, int32_t cut)
{
    std::vector<oosg::WidenRange> ranges;
    int64_t staged = 0;
    for (int32_t s = 0; s < p.S; ++s)
    {
        const int32_t q = p.parent[s];
        if (q < 0 || p.levelOf[s] >= cut || p.levelOf[q] < cut)
            continue;
        const int64_t values = static_cast<int64_t>(p.nrow[s]) * p.nrow[s];
        for (int64_t a = 0; a < values; a += 65536)
        {
            oosg::WidenRange r{};
            r.src = reinterpret_cast<const float*>(static_cast<uintptr_t>(p.updateOff[s] + a));     // offsets for now
            r.tmp = reinterpret_cast<float*>(static_cast<uintptr_t>(staged));
            r.dst = reinterpret_cast<double*>(static_cast<uintptr_t>(p.updateOff[s] + a));
            r.count = std::min<int64_t>(65536, values - a);
            staged += r.count;
            ranges.push_back(r);
        }
    }
    ensureCounted(p, p.widenTemp, static_cast<size_t>(std::max<int64_t>(staged, 1)) * sizeof(float),
                  "factor: the cut's update matrices");
    float* arena32 = p.updates.as<float>();
    double* arena64 = p.updates.as<double>();
    float* temp = p.widenTemp.as<float>();
    for (oosg::WidenRange& r : ranges)
    {
        r.src = arena32 + reinterpret_cast<uintptr_t>(r.src);
        r.tmp = temp + reinterpret_cast<uintptr_t>(r.tmp);
        r.dst = arena64 + reinterpret_cast<uintptr_t>(r.dst);
    }
    p.widenRanges.reset();
    if (!ranges.empty())
        p.widenRanges = upload(ranges, p.stream, "factor: the cut's update matrices");
    else
        p.widenRanges.bytes = 0;
}

// The numeric factorisation behind oosg_factorise (and the combined factorisation and solve).
void factoriseInto(oosg_plan* plan, const oosg_matrix* matrix, const oosg_factor_options* options, oosg_factor** out)
{
    {
        {
            if (plan == nullptr || matrix == nullptr || out == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "factorise: null argument");
            *out = nullptr;
            if (matrix->plan != plan)
                fail(OOSG_INVALID_ARGUMENT, "factorise: the matrix was made for another plan");
            oosg_factor_options opts{};
            opts.max_ratio = 1.0e7;
            if (options != nullptr)
                opts = *options;
            if (opts.pivoting != OOSG_POSITIVE_DEFINITE && opts.pivoting != OOSG_INDEFINITE)
                fail(OOSG_INVALID_ARGUMENT, "factorise: unknown pivoting mode");
            if (plan->trailing >= 0 && opts.pivoting != OOSG_POSITIVE_DEFINITE)
                // requirePositiveDefinite (src/linalg/supernodal/PartialFactor.cpp:24-30)
                fail(OOSG_INVALID_ARGUMENT,
                     "PartialFactor: the interior of a unit is factorised in positive definite mode (a delayed pivot "
                     "would move an interior equation into the boundary); pass PivotingMode::PositiveDefinite");
            if (opts.precision == OOSG_MIXED && (plan->outOfCore || plan->trailing >= 0 ||
                                                 opts.pivoting != OOSG_POSITIVE_DEFINITE))
                fail(OOSG_INVALID_ARGUMENT, "factorise: a mixed factor is positive definite, in core and whole");
            if (plan->outOfCore && opts.pivoting != OOSG_POSITIVE_DEFINITE)
                fail(OOSG_INVALID_ARGUMENT, "factorise: an out-of-core plan factorises in positive definite mode");
            if (plan->trailing >= 0 && opts.precision != OOSG_FP64)
                fail(OOSG_INVALID_ARGUMENT, "PartialFactor: the interior factor and K* are computed in FP64; pass "
                                            "Precision.FP64");
            if (opts.pivoting == OOSG_INDEFINITE && opts.precision != OOSG_FP64)
                fail(OOSG_INVALID_ARGUMENT, "factorise: indefinite factors are FP64 (their inertia decides Sturm "
                                            "counts)");
            if (opts.pivoting == OOSG_INDEFINITE)
            {
                setDevice(plan);
                auto f = std::make_unique<oosg_factor>();
                f->plan = plan;
                f->precision = OOSG_FP64;
                const hipStream_t st = plan->stream;
                check(hipEventRecord(plan->start, st), "hipEventRecord");
                factorisePivoted(*plan, *matrix, *f);
                check(hipEventRecord(plan->stop, st), "hipEventRecord");
                check(hipStreamSynchronize(st), "factor");
                float ms = 0.0f;
                check(hipEventElapsedTime(&ms, plan->start, plan->stop), "hipEventElapsedTime");
                f->info.factor_seconds = ms * 1e-3;
                f->info.flops = plan->flops;
                f->info.precision = OOSG_FP64;
                finishPivoted(*f);
                holdPlan(plan);
                *out = f.release();
                return;
            }
            if (opts.precision != OOSG_FP64 && opts.precision != OOSG_FP32 && opts.precision != OOSG_MIXED)
                fail(OOSG_INVALID_ARGUMENT, "factorise: unknown precision");
            setDevice(plan);
            pollCancellation(*plan);
            auto f = std::make_unique<oosg_factor>();
            f->plan = plan;
            f->precision = opts.precision;
            {
                FactorStorage storage = takeStorage(*plan);
                f->panels = std::move(storage.panels);
                f->linv = std::move(storage.linv);
                f->dDiag = std::move(storage.diag);
                f->scale = std::move(storage.scale);
                f->panels32 = std::move(storage.panels32);
                f->linv32 = std::move(storage.linv32);
            }
            if (plan->outOfCore)
            {
                const size_t values = static_cast<size_t>(plan->spoolPanelBase[plan->levels] +
                                                          plan->spoolLinvBase[plan->levels]);
                f->spoolBytes = values * valueBytes(opts.precision);
                check(hipHostMalloc(&f->spool, std::max<size_t>(f->spoolBytes, 8), hipHostMallocDefault),
                      "factor: the spool (pinned host memory)");
                f->traffic.spooled_bytes = f->spoolBytes;
                f->traffic.written_bytes = f->spoolBytes;
            }
            const bool mixed = opts.precision == OOSG_MIXED;
            if (mixed)
            {
                f->fp32Levels = std::clamp(opts.fp32_levels, 0, plan->levels);
                // Kept with the factor storage for the next mixed factor (no allocation once one was made).
                ensureCounted(*plan, f->panels32,
                              static_cast<size_t>(std::max<int64_t>(plan->panelOff[plan->S], 1)) * sizeof(float),
                              "factor: FP32 panels (mixed)");
                ensureCounted(*plan, f->linv32,
                              static_cast<size_t>(std::max<int64_t>(plan->linvOff[plan->S], 1)) * sizeof(float),
                              "factor: FP32 inverses (mixed)");
                prepareWidening(*plan, f->fp32Levels);
            }
            DeviceBuffer& diag = f->dDiag;
            DeviceBuffer& zero = plan->pivotStatus;
            const hipStream_t st = plan->stream;
            check(hipEventRecord(plan->start, st), "hipEventRecord");
            const auto body = [&](bool cancellable)
            {
                check(oosg::launchFillInt32(zero.as<int32_t>(), plan->S, INT32_MAX, st), "factor: status");
                if (mixed)
                {
                    const int32_t cut = f->fp32Levels;
                    runFactor<float>(*plan, *matrix, f->panels32.as<float>(), f->linv32.as<float>(),
                                     diag.as<double>(), zero.as<int32_t>(), true, cancellable, nullptr, 0, cut);
                    check(oosg::launchWidenRanges(plan->widenRanges.as<oosg::WidenRange>(),
                                                  static_cast<int>(plan->widenRanges.bytes / sizeof(oosg::WidenRange)),
                                                  st),
                          "factor: the cut's update matrices");
                    runFactor<double>(*plan, *matrix, f->panels.as<double>(), f->linv.as<double>(), diag.as<double>(),
                                      zero.as<int32_t>(), true, cancellable, nullptr, cut);
                }
                else if (opts.precision == OOSG_FP32)
                    runFactor<float>(*plan, *matrix, f->panels.as<float>(), f->linv.as<float>(), diag.as<double>(),
                                     zero.as<int32_t>(), true, cancellable, f->spool);
                else
                    runFactor<double>(*plan, *matrix, f->panels.as<double>(), f->linv.as<double>(),
                                      diag.as<double>(), zero.as<int32_t>(), false, cancellable, f->spool);
            };
            if (plan->graphs && !plan->profiler.enabled && !mixed && !plan->eager[opts.precision] && !plan->outOfCore)
                plan->factorGraph[opts.precision].run(
                    st,
                    {f->panels.ptr, f->linv.ptr, diag.ptr, zero.ptr, matrix->dValues.ptr, matrix->dPosition.ptr,
                     matrix->dColumn.ptr, matrix->dOffsets.ptr, matrix->dRows.ptr, matrix->dScale.ptr,
                     plan->updates.ptr, reinterpret_cast<const void*>(static_cast<intptr_t>(plan->concurrent[opts.precision]))},
                    [&] { body(false); });
            else
                body(true);
            if (plan->trailing >= 0)
                // K*: the trailing front assembled once every level is done (MultifrontalDriver::assembleTrailing,
                // MultifrontalDriver.cpp:567-599), into the trailing supernode's panel (its lower triangle).
                check(oosg::launchAssembly<double>(
                          plan->dp,
                          oosg::DeviceMatrix{matrix->dOffsets.as<int64_t>(), matrix->dRows.as<int32_t>(),
                                             matrix->dValues.as<double>(), matrix->dPosition.as<int32_t>(),
                                             matrix->dColumn.as<int32_t>(), nullptr},
                          f->panels.as<double>(), plan->updates.as<double>(),
                          plan->dTrailingBlocks.as<oosg::AssemblyBlock>(),
                          plan->dTrailingItems.as<oosg::AssemblyItem>(),
                          static_cast<int>(plan->trailingBlocks.size()), st),
                      "factor: the trailing front (K*)");
            check(hipEventRecord(plan->stop, st), "hipEventRecord");
            if (plan->aheadLevel > 0)
                check(hipStreamWaitEvent(st, plan->aheadEvents[LANES], 0), "factor: sweep ahead"); // its join
            if (scaledPrecision(opts.precision))
                // The equations' scales the pivots are divided by, kept with the factor.
                check(hipMemcpyAsync(f->scale.ptr, matrix->dScale.ptr, static_cast<size_t>(plan->n) * sizeof(double),
                                     hipMemcpyDeviceToDevice, st),
                      "factor: scales");
            // The ratio test's statistics over the pivots, on the device.
            std::vector<double> referenceHost;
            DeviceBuffer reference;
            if (opts.reference_diagonal != nullptr)
            {
                referenceHost.assign(opts.reference_diagonal, opts.reference_diagonal + plan->n);
                reference = upload(referenceHost, st, "factor: reference diagonal");
            }
            check(oosg::launchPivotStatistics(diag.as<double>(),
                                              scaledPrecision(opts.precision) ? f->scale.as<double>() : nullptr,
                                              matrix->dDiagonal.as<double>(),
                                              reference.ptr != nullptr ? reference.as<double>() : nullptr,
                                              plan->eliminated,
                                              opts.max_ratio > 0 ? opts.max_ratio : 1.0e7,
                                              plan->pivotStats.as<oosg::PivotStatistics>(), st),
                  "factor: pivot statistics");
            check(oosg::launchSupernodeGrowth(diag.as<double>(),
                                              scaledPrecision(opts.precision) ? f->scale.as<double>() : nullptr,
                                              matrix->dDiagonal.as<double>(), plan->dp.first, plan->dp.ncol, plan->S,
                                              plan->eliminated, plan->supernodeGrowth.as<float>(), st),
                  "factor: pivot growth");
            plan->growthRecorded = true;
            std::vector<int32_t> zeroAt(static_cast<size_t>(plan->S));
            oosg::PivotStatistics stats{};
            check(hipMemcpyAsync(zeroAt.data(), zero.ptr, zeroAt.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st),
                  "factor: status");
            check(hipMemcpyAsync(&stats, plan->pivotStats.ptr, sizeof(stats), hipMemcpyDeviceToHost, st),
                  "factor: pivot statistics");
            check(hipStreamSynchronize(st), "factor");
            plan->profiler.collect();
            pollCancellation(*plan);
            float ms = 0.0f;
            check(hipEventElapsedTime(&ms, plan->start, plan->stop), "hipEventElapsedTime");
            f->info.factor_seconds = ms * 1e-3;
            f->info.flops = plan->flops;
            f->info.precision = opts.precision;
            f->info.fp32_levels = f->fp32Levels;
            // FP32: the pivots of S·A·S divided by s_j² are those of A (same signs: the inertia is kept).
            finish(*plan, *matrix, *f, zeroAt, stats, opts);
            holdPlan(plan);
            *out = f.release();
        }
    }
}

} // namespace

extern "C" {

void oosg_factor_destroy(oosg_factor* f)
{
    if (f == nullptr)
        return;
    oosg_plan* plan = f->plan;
    (void)hipSetDevice(plan->device);
    delete f; // its buffers go back to the plan's caches
    releasePlan(plan);
    (void)hipGetLastError();
}

} // extern "C"

namespace
{

// A partial factor's tiles of 16 loads through the plan's vectors: each column block of `loads` (host, n rows) is
// uploaded, laid out row-major in p.solveX and swept forward over the interior; `after` then works on p.solveX.
template <typename After>
void partialTiles(const oosg_factor& f, const double* loads, int64_t ld, int64_t nrhs, After&& after)
{
    oosg_plan& p = *f.plan;
    setDevice(&p);
    ensureSolveBuffers(p);
    double* staging = p.cgB.as<double>();
    for (int64_t c0 = 0; c0 < nrhs; c0 += 16)
    {
        if (c0 > 0)
            pollCancellation(p);
        const int k = static_cast<int>(std::min<int64_t>(16, nrhs - c0));
        check(hipMemcpy2DAsync(staging, p.n * sizeof(double), loads + c0 * ld, ld * sizeof(double),
                               p.n * sizeof(double), k, hipMemcpyHostToDevice, p.stream),
              "partial factor: upload");
        check(oosg::launchToRowMajor(staging, p.n, p.solveX.as<double>(), p.n, k, nullptr, p.stream),
              "partial factor: layout");
        forwardSweep<double>(p, f.panels.as<double>(), f.linv.as<double>(), k, 0, p.levels, p.stream);
        after(c0, k, staging);
        check(hipStreamSynchronize(p.stream), "partial factor");
    }
    p.profiler.collect();
}

const oosg_plan& partialPlan(const oosg_factor* f, const char* operation)
{
    if (f == nullptr)
        fail(OOSG_INVALID_ARGUMENT, std::string("PartialFactor::") + operation + ": null factor");
    if (f->plan->trailing < 0)
        fail(OOSG_INVALID_ARGUMENT, std::string("PartialFactor::") + operation +
                                        ": the factor is not a partial one (make its plan with "
                                        "oosg_plan_create_partial)");
    return *f->plan;
}

} // namespace

extern "C" {

oosg_status oosg_partial_schur_complement(const oosg_factor* f, double* out, int64_t ld)
{
    return guarded(
        [&]
        {
            const oosg_plan& p = partialPlan(f, "schurComplement");
            const int64_t t = p.trailingDofs;
            if (out == nullptr || ld < std::max<int64_t>(t, 1))
                fail(OOSG_INVALID_ARGUMENT, "PartialFactor::schurComplement: null output or leading dimension below "
                                            "the trailing count");
            setDevice(&p);
            const double* kstar = f->panels.as<double>() + p.panelOff[p.trailing];
            check(hipMemcpy2DAsync(out, ld * sizeof(double), kstar, t * sizeof(double), t * sizeof(double), t,
                                   hipMemcpyDeviceToHost, p.stream),
                  "PartialFactor::schurComplement");
            check(hipStreamSynchronize(p.stream), "PartialFactor::schurComplement");
            // The lower triangle was assembled: both triangles handed over (MultifrontalDriver.cpp:593-596).
            for (int64_t j = 0; j < t; ++j)
                for (int64_t i = j + 1; i < t; ++i)
                    out[j + i * ld] = out[i + j * ld];
        });
}

oosg_status oosg_partial_condense_loads(const oosg_factor* f, const double* loads, int64_t ld, int64_t nrhs,
                                        double* condensed, int64_t ldc)
{
    return guarded(
        [&]
        {
            const oosg_plan& p = partialPlan(f, "condenseLoads");
            const int64_t t = p.trailingDofs;
            if ((loads == nullptr || condensed == nullptr) && nrhs > 0)
                fail(OOSG_INVALID_ARGUMENT, "PartialFactor::condenseLoads: null argument");
            if (ld < std::max<int64_t>(p.n, 1) || ldc < std::max<int64_t>(t, 1) || nrhs < 0)
                fail(OOSG_INVALID_ARGUMENT, "PartialFactor::condenseLoads: a leading dimension below the row count");
            if (nrhs == 0)
                return;
            oosg_plan& plan = *f->plan;
            const int32_t T = plan.trailing;
            partialTiles(*f, loads, ld, nrhs,
                         [&](int64_t c0, int k, double* staging)
                         {
                             double* X = plan.solveX.as<double>();
                             check(oosg::launchTrailingGather(X, plan.solveU.as<double>(), k, plan.first[T],
                                                              plan.ncol[T], plan.dTrailingPullStart.as<int32_t>(),
                                                              plan.dTrailingPulls.as<int32_t>(), plan.stream),
                                   "PartialFactor::condenseLoads");
                             check(oosg::launchFromRowMajor(X, staging, plan.n, plan.n, k, nullptr, 0, plan.stream),
                                   "PartialFactor::condenseLoads");
                             check(hipMemcpy2DAsync(condensed + c0 * ldc, ldc * sizeof(double),
                                                    staging + plan.eliminated, plan.n * sizeof(double),
                                                    t * sizeof(double), k, hipMemcpyDeviceToHost, plan.stream),
                                   "PartialFactor::condenseLoads: download");
                         });
        });
}

oosg_status oosg_partial_recover(const oosg_factor* f, double* loads_to_solution, int64_t ld, int64_t nrhs,
                                 const double* trailing_values, int64_t ldt)
{
    return guarded(
        [&]
        {
            const oosg_plan& p = partialPlan(f, "recover");
            const int64_t t = p.trailingDofs;
            if ((loads_to_solution == nullptr || (trailing_values == nullptr && t > 0)) && nrhs > 0)
                fail(OOSG_INVALID_ARGUMENT, "PartialFactor::recover: null argument");
            if (ld < std::max<int64_t>(p.n, 1) || ldt < std::max<int64_t>(t, 1) || nrhs < 0)
                fail(OOSG_INVALID_ARGUMENT, "PartialFactor::recover: a leading dimension below the row count");
            if (nrhs == 0)
                return;
            oosg_plan& plan = *f->plan;
            partialTiles(*f, loads_to_solution, ld, nrhs,
                         [&](int64_t c0, int k, double* staging)
                         {
                             double* X = plan.solveX.as<double>();
                             // The trailing rows (f* after the forward sweep) replaced by u_b, then D⁻¹ and the
                             // backward sweep over the interior.
                             check(hipMemcpy2DAsync(staging + plan.eliminated, plan.n * sizeof(double),
                                                    trailing_values + c0 * ldt, ldt * sizeof(double),
                                                    t * sizeof(double), k, hipMemcpyHostToDevice, plan.stream),
                                   "PartialFactor::recover: upload");
                             check(oosg::launchToRowMajor(staging + plan.eliminated, plan.n,
                                                          X + static_cast<int64_t>(plan.eliminated) * k, t, k,
                                                          nullptr, plan.stream),
                                   "PartialFactor::recover: layout");
                             check(oosg::launchDiagonalSolve(X, f->dDiag.as<double>(), plan.eliminated, k,
                                                             plan.stream),
                                   "PartialFactor::recover: diagonal");
                             backwardSweep<double>(plan, f->panels.as<double>(), f->linv.as<double>(), k);
                             check(oosg::launchFromRowMajor(X, staging, plan.n, plan.n, k, nullptr, 0, plan.stream),
                                   "PartialFactor::recover: layout");
                             check(hipMemcpy2DAsync(loads_to_solution + c0 * ld, ld * sizeof(double), staging,
                                                    plan.n * sizeof(double), plan.n * sizeof(double), k,
                                                    hipMemcpyDeviceToHost, plan.stream),
                                   "PartialFactor::recover: download");
                         });
        });
}

oosg_status oosg_factor_out_of_core_traffic(const oosg_factor* f, oosg_out_of_core_traffic* traffic)
{
    return guarded(
        [&]
        {
            if (f == nullptr || traffic == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "out of core: null argument");
            *traffic = f->traffic;
        });
}

oosg_status oosg_factor_issues(const oosg_factor* f, oosg_pivot_issue* issues, int64_t capacity, int64_t* recorded)
{
    return guarded(
        [&]
        {
            if (f == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "issues: null factor");
            if (recorded != nullptr)
                *recorded = static_cast<int64_t>(f->issues.size());
            if (issues != nullptr && capacity > 0)
                std::copy(f->issues.begin(),
                          f->issues.begin() + std::min<int64_t>(capacity, static_cast<int64_t>(f->issues.size())),
                          issues);
        });
}

oosg_status oosg_factor_info_get(const oosg_factor* f, oosg_factor_info* info)
{
    return guarded(
        [&]
        {
            if (f == nullptr || info == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "factor info: null argument");
            *info = f->info;
        });
}

oosg_status oosg_factor_panel(const oosg_factor* f, int32_t s, double* values, int32_t* rows, int32_t* eliminated)
{
    return guarded(
        [&]
        {
            if (f == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "panel: null factor");
            const oosg_plan& p = *f->plan;
            if (s < 0 || s >= p.S)
                fail(OOSG_INVALID_ARGUMENT, "panel: no supernode " + std::to_string(s));
            if (f->pivoted)
            {
                const PivotedFactor& pf = *f->pivoted;
                const int32_t e = pf.eliminated[s], N = pf.order[s];
                if (eliminated != nullptr)
                    *eliminated = e;
                if (rows != nullptr)
                    std::copy(pf.panelRows.begin() + pf.rowsStart[s], pf.panelRows.begin() + pf.rowsStart[s + 1], rows);
                if (values != nullptr && e > 0)
                {
                    setDevice(&p);
                    check(hipMemcpy(values, f->panels.as<double>() + pf.structure->panelOff[s],
                                    static_cast<size_t>(N) * e * sizeof(double), hipMemcpyDeviceToHost),
                          "panel: copy");
                }
                return;
            }
            const int32_t c = p.ncol[s], r = p.nrow[s];
            const size_t count = static_cast<size_t>(c) * (c + r);
            if (eliminated != nullptr)
                *eliminated = c;
            if (rows != nullptr)
            {
                for (int32_t k = 0; k < c; ++k)
                    rows[k] = p.first[s] + k;
                std::copy(p.rows.begin() + p.rowStart[s], p.rows.begin() + p.rowStart[s + 1], rows + c);
            }
            if (values == nullptr)
                return;
            setDevice(&p);
            if (p.outOfCore)
            {
                // From the spool: the level's panels, the supernode's within them.
                const int64_t at = p.spoolPanelBase[p.levelOf[s]] + p.panelOff[s];
                if (f->precision == OOSG_FP32)
                    std::copy_n(static_cast<const float*>(f->spool) + at, count, values);
                else
                    std::copy_n(static_cast<const double*>(f->spool) + at, count, values);
            }
            else if (f->precision == OOSG_FP32 ||
                     (f->precision == OOSG_MIXED && p.levelOf[s] >= 0 && p.levelOf[s] < f->fp32Levels))
            {
                std::vector<float> tmp(count);
                const float* from = (f->precision == OOSG_MIXED ? f->panels32 : f->panels).as<float>();
                check(hipMemcpy(tmp.data(), from + p.panelOff[s], count * sizeof(float), hipMemcpyDeviceToHost),
                      "panel: copy");
                std::copy(tmp.begin(), tmp.end(), values);
            }
            else
                check(hipMemcpy(values, f->panels.as<double>() + p.panelOff[s], count * sizeof(double),
                                hipMemcpyDeviceToHost),
                      "panel: copy");
            // The device never reads or writes the values above the diagonal (their memory may hold other values, an
            // out-of-core window another level's): they are zeros here, as in the reference's panels.
            const int64_t n = c + r;
            for (int64_t j = 1; j < c; ++j)
                std::fill(values + j * n, values + j * n + j, 0.0);
        });
}

oosg_status oosg_factor_panel_shape(const oosg_factor* f, int32_t s, int32_t* row_count, int32_t* eliminated)
{
    return guarded(
        [&]
        {
            if (f == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "panel: null factor");
            const oosg_plan& p = *f->plan;
            if (s < 0 || s >= p.S)
                fail(OOSG_INVALID_ARGUMENT, "panel: no supernode " + std::to_string(s));
            if (row_count != nullptr)
                *row_count = f->pivoted ? f->pivoted->order[s] : p.ncol[s] + p.nrow[s];
            if (eliminated != nullptr)
                *eliminated = f->pivoted ? f->pivoted->eliminated[s] : p.ncol[s];
        });
}

oosg_status oosg_factor_subdiagonal(const oosg_factor* f, int32_t s, double* subdiagonal)
{
    return guarded(
        [&]
        {
            if (f == nullptr || subdiagonal == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "panel: null argument");
            const oosg_plan& p = *f->plan;
            if (s < 0 || s >= p.S)
                fail(OOSG_INVALID_ARGUMENT, "panel: no supernode " + std::to_string(s));
            if (!f->pivoted)
            {
                std::fill(subdiagonal, subdiagonal + p.ncol[s], 0.0);
                return;
            }
            const PivotedFactor& pf = *f->pivoted;
            const int32_t first = pf.structure->first[s];
            std::copy(pf.sub_.begin() + first, pf.sub_.begin() + first + pf.eliminated[s], subdiagonal);
        });
}

oosg_status oosg_factor_diagonal(const oosg_factor* f, double* d)
{
    return guarded(
        [&]
        {
            if (f == nullptr || d == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "diagonal: null argument");
            setDevice(f->plan);
            const std::vector<double>& diag = f->pivoted ? f->diag : hostDiagonal(*f);
            std::copy(diag.begin(), diag.end(), d);
        });
}

oosg_status oosg_solve_device(const oosg_factor* f, double* b, int64_t ld, int64_t nrhs)
{
    return guarded(
        [&]
        {
            if (f == nullptr || (b == nullptr && nrhs > 0))
                fail(OOSG_INVALID_ARGUMENT, "FactorSolve::solve: null argument");
            if (ld < std::max<int64_t>(f->plan->n, 1) || nrhs < 0)
                fail(OOSG_INVALID_ARGUMENT, "FactorSolve::solve: the leading dimension is below the row count");
            if (nrhs == 0 || f->plan->n == 0)
                return;
            setDevice(f->plan);
            followCaller(*f->plan);
            solveDevice(*f, b, ld, nrhs);
            check(hipStreamSynchronize(f->plan->stream), "solve");
            f->plan->profiler.collect();
        });
}

oosg_status oosg_solve(const oosg_factor* f, double* b, int64_t ld, int64_t nrhs)
{
    return guarded(
        [&]
        {
            if (f == nullptr || (b == nullptr && nrhs > 0))
                fail(OOSG_INVALID_ARGUMENT, "FactorSolve::solve: null argument");
            oosg_plan& p = *f->plan;
            if (ld < std::max<int64_t>(p.n, 1) || nrhs < 0)
                fail(OOSG_INVALID_ARGUMENT, "FactorSolve::solve: the leading dimension is below the row count");
            if (nrhs == 0 || p.n == 0)
                return;
            setDevice(&p);
            // Tiles of 16 columns through the plan's refinement vector (column-major n × 16 there): no allocation.
            double* staging = p.cgB.as<double>();
            for (int64_t c0 = 0; c0 < nrhs; c0 += 16)
            {
                if (c0 > 0)
                    pollCancellation(p);
                const int64_t k = std::min<int64_t>(16, nrhs - c0);
                check(hipMemcpy2DAsync(staging, p.n * sizeof(double), b + c0 * ld, ld * sizeof(double),
                                       p.n * sizeof(double), k, hipMemcpyHostToDevice, p.stream),
                      "solve: upload");
                solveDevice(*f, staging, p.n, k);
                check(hipMemcpy2DAsync(b + c0 * ld, ld * sizeof(double), staging, p.n * sizeof(double),
                                       p.n * sizeof(double), k, hipMemcpyDeviceToHost, p.stream),
                      "solve: download");
                check(hipStreamSynchronize(p.stream), "solve");
            }
            p.profiler.collect();
        });
}

oosg_status oosg_matrix_multiply_device(const oosg_matrix* m, const double* x, double* y, int64_t ld, int64_t nrhs)
{
    return guarded(
        [&]
        {
            if (m == nullptr || x == nullptr || y == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "SymmetricCsc::multiply: null argument");
            if (ld < std::max<int64_t>(m->n, 1) || nrhs < 0)
                fail(OOSG_INVALID_ARGUMENT, "SymmetricCsc::multiply: the leading dimension is below the row count");
            setDevice(m->plan);
            followCaller(*m->plan);
            check(oosg::launchResidual(m->dCsrOffsets.as<int64_t>(), m->dCsrCols.as<int32_t>(),
                                       m->dCsrValues.as<double>(), x, nullptr, y, ld, m->n, static_cast<int>(nrhs),
                                       m->plan->stream),
                  "multiply");
            check(hipStreamSynchronize(m->plan->stream), "multiply");
        });
}

namespace
{

// The refined solve of columns [c0, nrhs) by tiles of 16 (the first tile's first forward sweep possibly done below
// `from`), the report merged over the tiles: the most solves, the largest backward error, all converged.
void refineTiles(const oosg_factor& f, const oosg_matrix& m, const double* b, double* x, int64_t ld, int64_t nrhs,
                 double target, int32_t maxSteps, oosg_solve_report& report, int32_t from)
{
    oosg_plan& p = *f.plan;
    report = {};
    report.converged = 1;
    for (int64_t c0 = 0; c0 < nrhs; c0 += 16)
    {
        if (c0 > 0)
            pollCancellation(p);
        const int k = static_cast<int>(std::min<int64_t>(16, nrhs - c0));
        oosg_solve_report tile{};
        if (scaledPrecision(f.precision))
            solveBlockPcg(f, m, b + c0 * ld, x + c0 * ld, ld, k, target, std::max(1, maxSteps), tile,
                          c0 == 0 ? from : 0);
        else
            solveRefinement(f, m, b + c0 * ld, x + c0 * ld, ld, k, target, std::max(1, maxSteps), tile,
                            c0 == 0 ? from : 0);
        report.steps = std::max(report.steps, tile.steps);
        report.backward_error = std::max(report.backward_error, tile.backward_error);
        report.converged = report.converged && tile.converged;
    }
}

} // namespace

oosg_status oosg_solve_refined_device(const oosg_factor* f, const oosg_matrix* m, const double* b, double* x,
                                      int64_t ld, int64_t nrhs, double target, int32_t max_steps,
                                      oosg_solve_report* report)
{
    oosg_solve_report local{};
    const oosg_status status = guarded(
        [&]
        {
            if (f == nullptr || m == nullptr || b == nullptr || x == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "refined solve: null argument");
            oosg_plan& p = *f->plan;
            if (m->plan != &p)
                fail(OOSG_INVALID_ARGUMENT, "refined solve: the matrix was made for another plan");
            if (ld < std::max<int64_t>(p.n, 1) || nrhs < 1)
                fail(OOSG_INVALID_ARGUMENT, "refined solve: at least one right-hand side, leading dimension >= n");
            setDevice(&p);
            followCaller(p);
            const hipStream_t st = p.stream;
            const auto t0 = std::chrono::steady_clock::now();
            refineTiles(*f, *m, b, x, ld, nrhs, target, max_steps, local, 0);
            check(hipStreamSynchronize(st), "refined solve");
            p.profiler.collect();
            local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (!local.converged)
                fail(OOSG_NOT_CONVERGED, "refined solve: the backward error " + std::to_string(local.backward_error) +
                                             " did not reach " + std::to_string(target) + " in " +
                                             std::to_string(local.steps) + " solves");
        });
    if (report != nullptr)
        *report = local;
    return status;
}

oosg_status oosg_solve_refined_device_noalloc(const oosg_factor* f, const oosg_matrix* m, const double* b, double* x,
                                              int64_t ld, int64_t nrhs, double target, int32_t max_steps,
                                              oosg_solve_report* report)
{
    oosg_solve_report local{};
    const oosg_status status = guarded(
        [&]
        {
            if (f == nullptr || m == nullptr || b == nullptr || x == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "refined solve: null argument");
            oosg_plan& p = *f->plan;
            if (m->plan != &p)
                fail(OOSG_INVALID_ARGUMENT, "refined solve: the matrix was made for another plan");
            if (ld < std::max<int64_t>(p.n, 1) || nrhs < 1)
                fail(OOSG_INVALID_ARGUMENT, "refined solve: at least one right-hand side, leading dimension >= n");
            if (p.scratch == nullptr)
                fail(OOSG_INTERNAL, "refined solve: no scratch array on this device");
            setDevice(&p);
            // The vectors the refined solve works in (sized for 16 right-hand sides, as the plan's own), laid out in
            // this device's scratch array: the plan's solve and refinement vectors, and a pivoted factor's sweep
            // vectors of its structure.
            oosg_plan* d = f->pivoted ? f->pivoted->structure.get() : nullptr;
            std::vector<DeviceBuffer*> targets = {&p.solveX, &p.solveU, &p.solvePartials, &p.solveNorms, &p.cgB,
                                                  &p.cgX,    &p.cgR,    &p.cgP,          &p.cgQ,         &p.cgPartial,
                                                  &p.cgDots};
            std::vector<uint64_t> sizes = solveBufferBytes(p);
            sizes.resize(targets.size());
            if (d != nullptr)
            {
                targets.insert(targets.end(), {&d->solveX, &d->solveU, &d->solvePartials});
                sizes.push_back(static_cast<uint64_t>(std::max(d->n, 1)) * 16 * sizeof(double));
                sizes.push_back(static_cast<uint64_t>(std::max<int64_t>(d->rowStart[d->S], 1)) * 16 * sizeof(double));
                sizes.push_back(static_cast<uint64_t>(std::max<int64_t>(d->partialSlots, 1)) * 64 * 16 *
                                sizeof(double));
            }
            std::vector<uint64_t> offsets(sizes.size());
            uint64_t total = 0;
            for (size_t k = 0; k < sizes.size(); ++k)
            {
                offsets[k] = total;
                total += (sizes[k] + 255) / 256 * 256;
            }
            if (total > oosg::SCRATCH_BYTES)
                fail(OOSG_INVALID_ARGUMENT,
                     "refined solve without allocation: the " + std::to_string(p.n) + " equations need " +
                         mebibytes(total) + " of scratch (16 right-hand-side columns), more than the " +
                         mebibytes(oosg::SCRATCH_BYTES) +
                         " of the library's scratch array; use oosg_solve_refined_device");
            // The views replace the plan's vectors for the call (launched one launch at a time: no graph capture, no
            // profiler events), and the plan's own come back at the end.
            struct Borrow
            {
                oosg_plan& p;
                std::vector<DeviceBuffer*> targets;
                std::vector<DeviceBuffer> saved;
                bool graphs, profiling;
                ~Borrow()
                {
                    for (size_t k = 0; k < targets.size(); ++k)
                        *targets[k] = std::move(saved[k]);
                    p.graphs = graphs;
                    p.profiler.enabled = profiling;
                }
            } borrow{p, targets, {}, p.graphs, p.profiler.enabled};
            unsigned char* base = static_cast<unsigned char*>(p.scratch);
            for (size_t k = 0; k < targets.size(); ++k)
            {
                borrow.saved.push_back(std::move(*targets[k]));
                *targets[k] = DeviceBuffer::view(base + offsets[k], sizes[k]);
            }
            p.graphs = false;
            p.profiler.enabled = false;
            followCaller(p);
            const hipStream_t st = p.stream;
            // The array keeps the last call's values: what is used is cleared first (a fill kernel).
            check(oosg::launchFillDouble(reinterpret_cast<double*>(base), static_cast<int64_t>(total / 8),
                                         static_cast<int64_t>(total / 8), 1, 0.0, st),
                  "refined solve: scratch");
            const auto t0 = std::chrono::steady_clock::now();
            refineTiles(*f, *m, b, x, ld, nrhs, target, max_steps, local, 0);
            check(hipStreamSynchronize(st), "refined solve");
            local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (!local.converged)
                fail(OOSG_NOT_CONVERGED, "refined solve: the backward error " + std::to_string(local.backward_error) +
                                             " did not reach " + std::to_string(target) + " in " +
                                             std::to_string(local.steps) + " solves");
        });
    if (report != nullptr)
        *report = local;
    return status;
}

oosg_status oosg_factorise_solve_device(oosg_plan* plan, const oosg_matrix* matrix, const oosg_factor_options* options,
                                        const double* b, double* x, int64_t ld, int64_t nrhs, double target,
                                        int32_t max_steps, oosg_factor** factor, oosg_solve_report* report)
{
    oosg_solve_report local{};
    const oosg_status status = guarded(
        [&]
        {
            if (plan == nullptr || matrix == nullptr || b == nullptr || x == nullptr || factor == nullptr)
                fail(OOSG_INVALID_ARGUMENT, "factorise and solve: null argument");
            *factor = nullptr;
            if (matrix->plan != plan)
                fail(OOSG_INVALID_ARGUMENT, "factorise and solve: the matrix was made for another plan");
            if (plan->trailing >= 0)
                fail(OOSG_INVALID_ARGUMENT, "factorise and solve: a partial plan's factor solves by "
                                            "oosg_partial_condense_loads and oosg_partial_recover");
            if (ld < std::max<int64_t>(plan->n, 1) || nrhs < 1)
                fail(OOSG_INVALID_ARGUMENT,
                     "factorise and solve: at least one right-hand side, leading dimension >= n");
            oosg_factor_options opts{};
            opts.max_ratio = 1.0e7;
            if (options != nullptr)
                opts = *options;
            const bool fp32 = opts.precision == OOSG_FP32;
            // The forward sweep goes ahead where the factorisation is launched one launch at a time on this call.
            const bool ahead = opts.pivoting == OOSG_POSITIVE_DEFINITE &&
                               (opts.precision == OOSG_FP64 || fp32) && plan->overlapLevel > 0 &&
                               !plan->profiler.enabled && (!plan->graphs || plan->eager[opts.precision]);
            setDevice(plan);
            followCaller(*plan);
            const hipStream_t st = plan->stream;
            const int k = static_cast<int>(std::min<int64_t>(nrhs, 16));
            const int64_t n = plan->n;
            if (ahead)
            {
                // The sweep's stream and events, made on first use.
                if (plan->sweepStream == nullptr)
                {
                    check(hipStreamCreateWithFlags(&plan->sweepStream, hipStreamNonBlocking), "hipStreamCreate");
                    for (hipEvent_t& e : plan->aheadEvents)
                        check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
                }
                // The first tile of b into the refinement's row-major copy, and scaled (FP32) into the sweeps'
                // vectors, with its column maxima in the first slot of the solve's norms: one pass.
                prepareRightHandSides(*plan, b, ld, plan->solveX.as<double>(),
                                      fp32 ? matrix->dScale.as<double>() : nullptr, k);
                plan->aheadLevel = plan->overlapLevel;
                plan->aheadK = k;
            }
            struct AheadOff
            {
                oosg_plan* p;
                ~AheadOff() { p->aheadLevel = -1; }
            } off{plan};
            factoriseInto(plan, matrix, &opts, factor);
            plan->aheadLevel = -1;
            const oosg_factor& f = **factor;
            const auto t0 = std::chrono::steady_clock::now();
            refineTiles(f, *matrix, b, x, ld, nrhs, target, max_steps, local, ahead ? plan->overlapLevel : 0);
            check(hipStreamSynchronize(st), "factorise and solve");
            plan->profiler.collect();
            local.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (!local.converged)
                fail(OOSG_NOT_CONVERGED, "refined solve: the backward error " + std::to_string(local.backward_error) +
                                             " did not reach " + std::to_string(target) + " in " +
                                             std::to_string(local.steps) + " solves");
        });
    if (report != nullptr)
        *report = local;
    return status;
}

// Diagnostic: one 16 × 16 × 4 MFMA product of each precision (A row-major 16 × 4, B row-major 4 × 16, D row-major
// 16 × 16), host arrays; checks the lane layout the tile kernel assumes.
oosg_status oosg_debug_mfma_probe(const double* a, const double* b, double* d, const float* af, const float* bf,
                                  float* df)
{
    return guarded(
        [&]
        {
            DeviceBuffer da, db, dd, fa, fb, fd;
            da.allocate(64 * 8, "probe");
            db.allocate(64 * 8, "probe");
            dd.allocate(256 * 8, "probe");
            fa.allocate(64 * 4, "probe");
            fb.allocate(64 * 4, "probe");
            fd.allocate(256 * 4, "probe");
            check(hipMemcpy(da.ptr, a, 64 * 8, hipMemcpyHostToDevice), "probe");
            check(hipMemcpy(db.ptr, b, 64 * 8, hipMemcpyHostToDevice), "probe");
            check(hipMemcpy(fa.ptr, af, 64 * 4, hipMemcpyHostToDevice), "probe");
            check(hipMemcpy(fb.ptr, bf, 64 * 4, hipMemcpyHostToDevice), "probe");
            check(oosg::launchMfmaProbe(da.as<double>(), db.as<double>(), dd.as<double>(), fa.as<float>(),
                                        fb.as<float>(), fd.as<float>(), nullptr),
                  "probe");
            check(hipDeviceSynchronize(), "probe");
            check(hipMemcpy(d, dd.ptr, 256 * 8, hipMemcpyDeviceToHost), "probe");
            check(hipMemcpy(df, fd.ptr, 256 * 4, hipMemcpyDeviceToHost), "probe");
        });
}

} // extern "C"
