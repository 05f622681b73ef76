// Kernel, KernelFactory and PreProcessor of the C++ layer.
#include "virtualfluids_amd.hpp"

namespace vfamd {

std::shared_ptr<K17CompressibleNavierStokes> K17CompressibleNavierStokes::getNewInstance(std::shared_ptr<Parameter> para,
                                                                                         int level)
{
    return std::shared_ptr<K17CompressibleNavierStokes>(new K17CompressibleNavierStokes(std::move(para), level));
}

// K17CompressibleNavierStokes::run (K17CompressibleNavierStokes.cu:83-101): every fluid node,
// CollisionTemplate::Default. A dense box whose solid nodes are masked (or that has none) takes
// the box kernel, which recomputes neighbours instead of loading them.
void K17CompressibleNavierStokes::run()
{
    auto p = para->getParD(level);
    const bool dense = p->nx > 0 && (uint64_t)p->nx * p->ny * p->nz == p->numberOfNodes;
    const bool all_covered = p->fluidMask != nullptr || p->numberOfFluidNodes == p->numberOfNodes;
    if (dense && all_covered)
        check(vf_k17_run_box(p.get(), VF_COLLISION_DEFAULT, p->stream));
    else
        check(vf_k17_run(p.get(), p->stream));
}

// K17CompressibleNavierStokes::runOnIndices (K17CompressibleNavierStokes.cu:103-145)
void K17CompressibleNavierStokes::runOnIndices(const unsigned int* indices, unsigned int size_indices,
                                               CollisionTemplate collisionTemplate, hipStream_t stream)
{
    auto p = para->getParD(level);
    check(vf_k17_run_on_indices(p.get(), indices, size_indices, (int32_t)collisionTemplate,
                                stream ? stream : p->stream));
}

// KernelFactoryImp::makeKernels / makeKernel (KernelFactoryImp.cpp:61-135)
std::vector<std::shared_ptr<Kernel>> KernelFactory::makeKernels(std::shared_ptr<Parameter> para)
{
    std::vector<std::shared_ptr<Kernel>> kernels;
    for (int level = 0; level <= para->getMaxLevel(); ++level)
        kernels.push_back(makeKernel(para, para->getMainKernel(), level));
    return kernels;
}

std::shared_ptr<Kernel> KernelFactory::makeKernel(std::shared_ptr<Parameter> para, const std::string& kernel, int level)
{
    if (kernel == collisionKernel::compressible::K17CompressibleNavierStokes) {
        switch (para->getTurbulenceModel()) {
        case TurbulenceModel::None: {
            auto k = K17CompressibleNavierStokes::getNewInstance(para, level);
            para->setKernelNeedsFluidNodeIndicesToRun(k->getKernelUsesFluidNodeIndices());
            return k;
        }
        case TurbulenceModel::Smagorinsky:
        case TurbulenceModel::AMD:
        case TurbulenceModel::QR:
            throw std::runtime_error("K17CompressibleNavierStokes with a turbulence model is not part of this build "
                                     "(TurbulenceModel::None only)");
        }
        throw std::runtime_error("Unknown turbulence model!");
    }
    throw std::runtime_error("KernelFactory does not know the KernelType.");
}

// InitNavierStokesCompressible::init (InitNavierStokesCompressible.cu:45-87)
void PreProcessor::init(int level)
{
    auto p = para->getParD(level);
    check(vf_InitNavierStokesCompressible(p.get(), p->stream));
}

} // namespace vfamd
