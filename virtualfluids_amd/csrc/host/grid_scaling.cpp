// GridScalingFactory and GridScalingKernelManager of the C++ layer
// (src/gpu/core/GridScaling/GridScalingFactory.cpp:38-69, GridScalingKernelManager.cpp:45-75,
//  GridScalingKernelManager.h:77-84).
#include "virtualfluids_amd.hpp"

namespace vfamd {

namespace {
using scale_launcher = int (*)(const vf_lbm_parameter*, const vf_lbm_parameter*, const vf_interpolation_cells*,
                               const vf_interpolation_cell_neighbor*, void*);

gridScaling wrap(scale_launcher fn, bool coarseToFine)
{
    return [fn, coarseToFine](LBMSimulationParameter* coarse, LBMSimulationParameter* fine, InterpolationCells* cells,
                              ICellNeigh& neigh, hipStream_t stream) {
        const bool offsets = coarseToFine ? coarse->hasOffsetsCoarseToFine : coarse->hasOffsetsFineToCoarse;
        check(fn(coarse, fine, cells, offsets ? &neigh : nullptr, stream ? stream : coarse->stream));
    };
}
} // namespace

gridScaling GridScalingFactory::getGridScalingFC(bool hasTurbulentViscosity) const
{
    if (gridScalingType != GridScaling::ScaleCompressible) return nullptr;
    if (hasTurbulentViscosity)
        throw std::runtime_error("scaleFineToCoarseCompressible with turbulent viscosity is not part of this build");
    return wrap(vf_ScaleFineToCoarseCompressible, false);
}

gridScaling GridScalingFactory::getGridScalingCF(bool hasTurbulentViscosity) const
{
    if (gridScalingType != GridScaling::ScaleCompressible) return nullptr;
    if (hasTurbulentViscosity)
        throw std::runtime_error("scaleCoarseToFineCompressible with turbulent viscosity is not part of this build");
    return wrap(vf_ScaleCoarseToFineCompressible, true);
}

namespace {
void checkScalingFunction(const gridScaling& fn, const InterpolationCells& cells, const std::string& name)
{
    if (!fn && cells.numberOfCells > 0) throw std::runtime_error("The scaling function " + name + " was not set!");
}
} // namespace

GridScalingKernelManager::GridScalingKernelManager(std::shared_ptr<Parameter> parameter,
                                                   GridScalingFactory* gridScalingFactory)
    : para(std::move(parameter))
{
    if (para->getMaxLevel() != 0) {
        if (!gridScalingFactory)
            throw std::runtime_error("There is more than one level, but no scalingFactory was provided.");
        const bool tv = para->getUseTurbulentViscosity();
        checkScalingFunction(gridScalingFactory->getGridScalingFC(tv), para->getParD(0)->fineToCoarse,
                             "scalingFineToCoarse");
        checkScalingFunction(gridScalingFactory->getGridScalingCF(tv), para->getParD(0)->coarseToFine,
                             "scalingCoarseToFine");
        scalingFineToCoarse = gridScalingFactory->getGridScalingFC(tv);
        scalingCoarseToFine = gridScalingFactory->getGridScalingCF(tv);
    }
}

void GridScalingKernelManager::runFineToCoarseKernelLB(int level, InterpolationCells* fineToCoarse,
                                                       ICellNeigh& neighborFineToCoarse, hipStream_t stream) const
{
    scalingFineToCoarse(para->getParD(level).get(), para->getParD(level + 1).get(), fineToCoarse, neighborFineToCoarse,
                        stream);
}

void GridScalingKernelManager::runCoarseToFineKernelLB(int level, InterpolationCells* coarseToFine,
                                                       ICellNeigh& neighborCoarseToFine, hipStream_t stream) const
{
    scalingCoarseToFine(para->getParD(level).get(), para->getParD(level + 1).get(), coarseToFine, neighborCoarseToFine,
                        stream);
}

} // namespace vfamd
