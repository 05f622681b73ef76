// BoundaryConditionFactory and BoundaryConditionKernelManager of the C++ layer
// (src/gpu/core/BoundaryConditions/BoundaryConditionFactory.cpp:49-217,
//  src/gpu/core/BoundaryConditions/BoundaryConditionKernelManager.cpp:52-344).
#include "virtualfluids_amd.hpp"

namespace vfamd {

namespace {

using launcher = int (*)(const vf_lbm_parameter*, const vf_q_boundary_conditions*, void*);

boundaryCondition wrap(launcher fn)
{
    return [fn](LBMSimulationParameter* p, QforBoundaryConditions* bc) { check(fn(p, bc, p->stream)); };
}

[[noreturn]] void not_on_path(const char* name)
{
    throw std::runtime_error(std::string(name) + " is not part of the K17 compressible path of this build");
}

} // namespace

boundaryCondition BoundaryConditionFactory::getVelocityBoundaryConditionPost(bool isGeometryBC) const
{
    const VelocityBC bc = isGeometryBC ? std::get<VelocityBC>(geometryBoundaryCondition) : velocityBoundaryCondition;
    switch (bc) {
    case VelocityBC::VelocityBounceBack: return wrap(vf_VelocityBounceBack);
    case VelocityBC::VelocityInterpolatedCompressible: return wrap(vf_VelocityInterpolatedCompressible);
    case VelocityBC::VelocityWithPressureInterpolatedCompressible:
        return wrap(vf_VelocityWithPressureInterpolatedCompressible);
    case VelocityBC::VelocityInterpolatedIncompressible: not_on_path("VelocityInterpolatedIncompressible");
    case VelocityBC::NotSpecified: break;
    }
    return nullptr;
}

boundaryCondition BoundaryConditionFactory::getNoSlipBoundaryConditionPost(bool isGeometryBC) const
{
    const NoSlipBC bc = isGeometryBC ? std::get<NoSlipBC>(geometryBoundaryCondition) : noSlipBoundaryCondition;
    switch (bc) {
    case NoSlipBC::NoSlipDelayBounceBack:
        // solid nodes never collide: the esoteric twist reflects populations by itself
        return [](LBMSimulationParameter*, QforBoundaryConditions*) {};
    case NoSlipBC::NoSlipBounceBack: return wrap(vf_NoSlipBounceBack);
    case NoSlipBC::NoSlipInterpolatedCompressible: return wrap(vf_NoSlipInterpolatedCompressible);
    case NoSlipBC::NoSlipInterpolatedIncompressible: not_on_path("NoSlipInterpolatedIncompressible");
    }
    return nullptr;
}

boundaryCondition BoundaryConditionFactory::getSlipBoundaryConditionPost(bool isGeometryBC) const
{
    const SlipBC bc = isGeometryBC ? std::get<SlipBC>(geometryBoundaryCondition) : slipBoundaryCondition;
    switch (bc) {
    case SlipBC::SlipCompressible: not_on_path("SlipCompressible");
    case SlipBC::SlipTurbulentViscosityCompressible: not_on_path("SlipTurbulentViscosityCompressible");
    case SlipBC::NotSpecified: break;
    }
    return nullptr;
}

boundaryCondition BoundaryConditionFactory::getGeometryBoundaryConditionPost() const
{
    if (std::holds_alternative<VelocityBC>(geometryBoundaryCondition)) return getVelocityBoundaryConditionPost(true);
    if (std::holds_alternative<NoSlipBC>(geometryBoundaryCondition)) return getNoSlipBoundaryConditionPost(true);
    return getSlipBoundaryConditionPost(true);
}

std::variant<boundaryCondition, boundaryConditionDirectional>
BoundaryConditionFactory::getPressureBoundaryConditionPre() const
{
    switch (pressureBoundaryCondition) {
    case PressureBC::PressureNonEquilibriumCompressible:
        return boundaryConditionDirectional([](LBMSimulationParameter* p, QforDirectionalBoundaryCondition* bc) {
            check(vf_PressureNonEquilibriumCompressible(p, bc, p->stream));
        });
    case PressureBC::PressureNonEquilibriumIncompressible: not_on_path("PressureNonEquilibriumIncompressible");
    case PressureBC::OutflowNonReflective: not_on_path("OutflowNonReflective");
    case PressureBC::OutflowNonReflectivePressureCorrection: not_on_path("OutflowNonReflectivePressureCorrection");
    case PressureBC::NotSpecified: break;
    }
    return boundaryCondition(nullptr);
}

bool BoundaryConditionFactory::hasDirectionalPressureBoundaryCondition() const
{
    return pressureBoundaryCondition != PressureBC::NotSpecified;
}

// ---- BoundaryConditionKernelManager ------------------------------------------------------------
namespace {
template <class F> void checkBoundaryCondition(const F& fn, uint32_t nodes, const std::string& name)
{
    if (!fn && nodes > 0) throw std::runtime_error("The boundary condition " + name + " was not set!");
}
} // namespace

BoundaryConditionKernelManager::BoundaryConditionKernelManager(std::shared_ptr<Parameter> parameter,
                                                               const BoundaryConditionFactory* bcFactory)
    : para(std::move(parameter))
{
    if (!bcFactory) throw std::runtime_error("BoundaryConditionKernelManager needs a BoundaryConditionFactory");
    velocityBoundaryConditionPost = bcFactory->getVelocityBoundaryConditionPost();
    noSlipBoundaryConditionPost = bcFactory->getNoSlipBoundaryConditionPost();
    slipBoundaryConditionPost = bcFactory->getSlipBoundaryConditionPost();
    geometryBoundaryConditionPost = bcFactory->getGeometryBoundaryConditionPost();
    if (bcFactory->hasDirectionalPressureBoundaryCondition())
        directionalPressureBoundaryConditionPre =
            std::get<boundaryConditionDirectional>(bcFactory->getPressureBoundaryConditionPre());
    else
        pressureBoundaryConditionPre = std::get<boundaryCondition>(bcFactory->getPressureBoundaryConditionPre());

    auto p = para->getParD(0);
    uint32_t directional = 0;
    for (const auto& bc : p->pressureBCDirectional) directional += bc.numberOfBCnodes > 0 ? 1 : 0;
    checkBoundaryCondition(velocityBoundaryConditionPost, p->velocityBC.numberOfBCnodes, "velocityBoundaryConditionPost");
    checkBoundaryCondition(noSlipBoundaryConditionPost, p->noSlipBC.numberOfBCnodes, "noSlipBoundaryConditionPost");
    checkBoundaryCondition(slipBoundaryConditionPost, p->slipBC.numberOfBCnodes, "slipBoundaryConditionPost");
    checkBoundaryCondition(geometryBoundaryConditionPost, p->geometryBC.numberOfBCnodes, "geometryBoundaryConditionPost");
    checkBoundaryCondition(pressureBoundaryConditionPre, p->pressureBC.numberOfBCnodes, "pressureBoundaryConditionPre");
    checkBoundaryCondition(directionalPressureBoundaryConditionPre, directional,
                           "directionalPressureBoundaryConditionPre");
}

void BoundaryConditionKernelManager::runVelocityBCKernelPost(int level) const
{
    auto p = para->getParD(level);
    if (p->velocityBC.numberOfBCnodes > 0) velocityBoundaryConditionPost(p.get(), &p->velocityBC);
}

void BoundaryConditionKernelManager::runNoSlipBCKernelPost(int level) const
{
    auto p = para->getParD(level);
    if (p->noSlipBC.numberOfBCnodes > 0) noSlipBoundaryConditionPost(p.get(), &p->noSlipBC);
}

void BoundaryConditionKernelManager::runSlipBCKernelPost(int level) const
{
    auto p = para->getParD(level);
    if (p->slipBC.numberOfBCnodes > 0) slipBoundaryConditionPost(p.get(), &p->slipBC);
}

void BoundaryConditionKernelManager::runGeometryBCKernelPost(int level) const
{
    auto p = para->getParD(level);
    if (p->geometryBC.numberOfBCnodes > 0) geometryBoundaryConditionPost(p.get(), &p->geometryBC);
}

// runPressureBCKernelPre (BoundaryConditionKernelManager.cpp:321-328)
void BoundaryConditionKernelManager::runPressureBCKernelPre(int level) const
{
    auto p = para->getParD(level);
    for (auto& bc : p->pressureBCDirectional)
        if (bc.numberOfBCnodes > 0) directionalPressureBoundaryConditionPre(p.get(), &bc);
    if (p->pressureBC.numberOfBCnodes > 0) pressureBoundaryConditionPre(p.get(), &p->pressureBC);
}

} // namespace vfamd
