// UpdateGrid27 of the C++ layer: the order of one time step on a level hierarchy
// (src/gpu/core/Calculation/UpdateGrid27.cpp:54-96, 228-308; refinement as
//  Refinement_noExchange, src/gpu/core/GridScaling/RefinementStrategy.cpp:169-178).
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

void UpdateGrid27::run(unsigned int steps)
{
    for (unsigned int i = 0; i < steps; ++i) updateGrid(para->getCoarse(), ++timestep);
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
