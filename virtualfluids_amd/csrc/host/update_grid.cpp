// UpdateGrid27 of the C++ layer: the order of one time step on a level hierarchy
// (src/gpu/core/Calculation/UpdateGrid27.cpp:54-96, 228-308; refinement as
//  Refinement_noExchange, src/gpu/core/GridScaling/RefinementStrategy.cpp:169-178).
#include <hip/hip_runtime.h>

#include "virtualfluids_amd.hpp"

namespace vfamd {

UpdateGrid27::UpdateGrid27(std::shared_ptr<Parameter> parameter, std::vector<std::shared_ptr<Kernel>> kernelList,
                           const BoundaryConditionFactory* bcFactory, GridScalingFactory* scalingFactory)
    : para(std::move(parameter)),
      kernels(std::move(kernelList)),
      bcKernelManager(para, bcFactory),
      gridScalingKernelManager(para, scalingFactory)
{
    if ((int)kernels.size() != para->getMaxLevel() + 1)
        throw std::runtime_error("UpdateGrid27 needs one kernel per level");
    for (int level = 1; level <= para->getMaxLevel(); ++level)
        if (!para->getParD(level)->isEvenTimestep)
            throw std::runtime_error("finer levels must start with the even parity (the interface kernels read "
                                     "and write them with it)");
    // a coarse step cycles through every level's populations; levels the caller left at 0 get
    // that sum (it picks the collision kernel: cache-resident or streaming)
    if (para->getMaxLevel() > 0) {
        uint64_t total = 0;
        for (int level = para->getCoarse(); level <= para->getFine(); ++level) {
            const auto d = para->getParD(level);
            total += 27ull * d->numberOfNodes * (d->storage == VF_STORAGE_FP32 ? 4u : 2u);
        }
        for (int level = para->getCoarse(); level <= para->getFine(); ++level)
            if (para->getParD(level)->workingSetBytes == 0) para->getParD(level)->workingSetBytes = total;
    }
}

void UpdateGrid27::updateGrid(int level, unsigned int t)
{
    if (level != para->getFine()) {
        updateGrid(level + 1, t);
        updateGrid(level + 1, t);
    }
    collision(level);
    postCollisionBC(level);
    swapBetweenEvenAndOddTimestep(level);
    preCollisionBC(level);
    if (level != para->getFine()) refinement(level);
}

namespace {

void hip_ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string("UpdateGrid27::run: ") + what + ": " + hipGetErrorString(e));
}

} // namespace

// With graphs, steps run in pairs from one HIP graph: after two coarse steps every level's
// parity is back where it started (the finer levels take four, eight, ... steps), so the launches
// of a pair, captured once with their arguments, are the launches of every pair. The first two
// steps run from the host (lazy set-up stays outside the capture), an odd last step too. The
// capture routes every level to one stream of this object; it waits for the levels' streams
// before the replays and they wait for it afterwards. run() returns once the replays are done
// (the graph is released there).
void UpdateGrid27::run(unsigned int steps)
{
    unsigned int eager = (useGraphs && steps >= 4) ? 2 : steps;
    for (unsigned int i = 0; i < eager; ++i) updateGrid(para->getCoarse(), ++timestep);
    steps -= eager;
    const unsigned int pairs = steps / 2;
    if (pairs > 0) {
        std::vector<hipStream_t> levelStreams;
        for (int level = para->getCoarse(); level <= para->getFine(); ++level)
            levelStreams.push_back(para->getParD(level)->stream);
        hipStream_t gs = nullptr;
        hip_ok(hipStreamCreateWithFlags(&gs, hipStreamNonBlocking), "stream");
        std::vector<hipEvent_t> events;
        auto join = [&](hipStream_t from, hipStream_t to) {
            hipEvent_t e = nullptr;
            hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
            events.push_back(e);
            hip_ok(hipEventRecord(e, from), "event record");
            hip_ok(hipStreamWaitEvent(to, e, 0), "stream wait");
        };
        for (hipStream_t s : levelStreams) join(s, gs);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hip_ok(hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal), "begin capture");
        for (int level = para->getCoarse(); level <= para->getFine(); ++level) para->getParD(level)->stream = gs;
        try {
            updateGrid(para->getCoarse(), timestep + 1);
            updateGrid(para->getCoarse(), timestep + 2);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(gs, &g);
            if (g) (void)hipGraphDestroy(g);
            for (int level = para->getCoarse(), i = 0; level <= para->getFine(); ++level, ++i)
                para->getParD(level)->stream = levelStreams[i];
            throw;
        }
        const hipError_t ec = hipStreamEndCapture(gs, &graph);
        for (int level = para->getCoarse(), i = 0; level <= para->getFine(); ++level, ++i)
            para->getParD(level)->stream = levelStreams[i];
        hip_ok(ec, "end capture");
        hip_ok(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
        for (unsigned int i = 0; i < pairs; ++i) hip_ok(hipGraphLaunch(exec, gs), "graph launch");
        timestep += 2 * pairs;
        for (hipStream_t s : levelStreams) join(gs, s);
        hip_ok(hipStreamSynchronize(gs), "synchronize");
        hip_ok(hipGraphExecDestroy(exec), "graph release");
        hip_ok(hipGraphDestroy(graph), "graph release");
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        hip_ok(hipStreamDestroy(gs), "stream release");
    }
    if (steps % 2) updateGrid(para->getCoarse(), ++timestep);
}

// collisionUsingIndices / collisionAllNodes (UpdateGrid27.cpp:98-119)
void UpdateGrid27::collision(int level)
{
    auto p = para->getParD(level);
    if (para->getKernelNeedsFluidNodeIndicesToRun() && p->fluidNodeIndices == nullptr && p->numberOfFluidNodes > 0)
        throw std::runtime_error("In collision: fluidNodeIndices or numberOfFluidNodes not defined");
    kernels.at(level)->run();
}

void UpdateGrid27::postCollisionBC(int level)
{
    bcKernelManager.runVelocityBCKernelPost(level);
    bcKernelManager.runNoSlipBCKernelPost(level);
    bcKernelManager.runSlipBCKernelPost(level);
    bcKernelManager.runGeometryBCKernelPost(level);
}

void UpdateGrid27::swapBetweenEvenAndOddTimestep(int level)
{
    auto p = para->getParD(level);
    p->isEvenTimestep = p->isEvenTimestep ? 0 : 1;
}

void UpdateGrid27::preCollisionBC(int level) { bcKernelManager.runPressureBCKernelPre(level); }

void UpdateGrid27::refinement(int level)
{
    auto p = para->getParD(level);
    if (p->fineToCoarse.numberOfCells > 0)
        gridScalingKernelManager.runFineToCoarseKernelLB(level, &p->fineToCoarse, p->neighborFineToCoarse, p->stream);
    if (p->coarseToFine.numberOfCells > 0)
        gridScalingKernelManager.runCoarseToFineKernelLB(level, &p->coarseToFine, p->neighborCoarseToFine, p->stream);
}

} // namespace vfamd
