/* C++ host layer of the gfx950 LBM time step.
 *
 * The classes the reference's GPU core exposes around its kernels, with the same names,
 * argument meaning and error behaviour, re-implemented over the C ABI of
 * virtualfluids_amd.h (every launch goes through one of its extern "C" entry points):
 *
 *   Kernel / K17CompressibleNavierStokes      src/gpu/core/Kernel/Kernel.h:44-52,
 *                                             src/gpu/core/Kernel/Compressible/NavierStokes/K17/K17CompressibleNavierStokes.cu:57-145
 *   KernelFactory                             src/gpu/core/Kernel/KernelFactory/KernelFactoryImp.cpp:61-135
 *   BoundaryConditionFactory                  src/gpu/core/BoundaryConditions/BoundaryConditionFactory.h:50-171
 *   BoundaryConditionKernelManager            src/gpu/core/BoundaryConditions/BoundaryConditionKernelManager.cpp:52-344
 *   GridScalingFactory / GridScalingKernelManager
 *                                             src/gpu/core/GridScaling/GridScalingFactory.h:42-66,
 *                                             src/gpu/core/GridScaling/GridScalingKernelManager.cpp:45-75
 *   PreProcessor (InitNavierStokesCompressible)
 *                                             src/gpu/core/PreProcessor/PreProcessorStrategy/InitNavierStokesCompressible/InitNavierStokesCompressible.cu:45-87
 *   UpdateGrid27                              src/gpu/core/Calculation/UpdateGrid27.cpp:54-96
 *
 * Errors: a failing entry point throws std::runtime_error carrying vf_get_last_error();
 * missing BC or scaling functions with nodes present throw at manager construction, as in
 * the reference. Launches are asynchronous on the level's stream (hipStream_t, default 0).
 */
#ifndef VIRTUALFLUIDS_AMD_HPP
#define VIRTUALFLUIDS_AMD_HPP

#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "virtualfluids_amd.h"

typedef struct ihipStream_t* hipStream_t;

namespace vfamd {

/* Throws std::runtime_error(vf_get_last_error()) for a non-zero status. */
void check(int status);

/* src/gpu/core/Calculation/Calculation.h:56-69 */
enum class CollisionTemplate : int32_t {
    Default = VF_COLLISION_DEFAULT,
    WriteMacroVars = VF_COLLISION_WRITE_MACRO_VARS,
    ApplyBodyForce = VF_COLLISION_APPLY_BODY_FORCE,
    AllFeatures = VF_COLLISION_ALL_FEATURES,
    SubDomainBorder = VF_COLLISION_SUBDOMAIN_BORDER
};

/* src/lbm/collision/TurbulentViscosity.h: only None is on this path */
enum class TurbulenceModel { Smagorinsky, AMD, QR, None };

/* src/gpu/core/PreProcessor/PreProcessorType.h */
enum class PreProcessorType { InitNavierStokesCompressible };

/* QforBoundaryConditions / QforDirectionalBoundaryCondition (Calculation.h:105-127) */
struct QforBoundaryConditions : vf_q_boundary_conditions {
    QforBoundaryConditions() : vf_q_boundary_conditions{} {}
};
struct QforDirectionalBoundaryCondition : vf_q_boundary_conditions {
    QforDirectionalBoundaryCondition() : vf_q_boundary_conditions{} {}
};
/* ICells / ICellNeigh (Calculation.h:72-89) */
struct InterpolationCells : vf_interpolation_cells {
    InterpolationCells() : vf_interpolation_cells{} {}
};
struct ICellNeigh : vf_interpolation_cell_neighbor {
    ICellNeigh() : vf_interpolation_cell_neighbor{} {}
};

/* The device view of one grid level (LBMSimulationParameter, Parameter.h:67-256): the C-ABI
 * level struct plus the BC lists, the interface to the next finer level and a stream. */
struct LBMSimulationParameter : vf_lbm_parameter {
    LBMSimulationParameter() : vf_lbm_parameter{} {}
    double viscosity = 0.0;
    QforBoundaryConditions velocityBC, noSlipBC, slipBC, geometryBC, pressureBC;
    std::vector<QforDirectionalBoundaryCondition> pressureBCDirectional;
    InterpolationCells fineToCoarse, coarseToFine; /* towards level + 1 */
    ICellNeigh neighborFineToCoarse, neighborCoarseToFine;
    bool hasOffsetsFineToCoarse = false, hasOffsetsCoarseToFine = false;
    hipStream_t stream = nullptr;
};

/* Simulation-wide settings and the level structs (the parts of Parameter, Parameter.h, that
 * the time step reads). Device pointers in the levels are owned by the caller (or by a
 * DeviceLevel below). */
class Parameter {
public:
    void setMaxLevel(int maxLevel);
    int getMaxLevel() const { return maxLevel; }
    int getCoarse() const { return 0; }
    int getFine() const { return maxLevel; }
    void setViscosityLB(double nu) { viscosity = nu; }
    double getViscosity() const { return viscosity; }
    void setQuadricLimiters(float p, float m, float d);
    void setForcesLB(float fx, float fy, float fz);
    void setStorage(vf_storage s) { storage = s; }
    void setTurbulenceModel(TurbulenceModel tm) { turbulenceModel = tm; }
    TurbulenceModel getTurbulenceModel() const { return turbulenceModel; }
    bool getUseTurbulentViscosity() const { return turbulenceModel != TurbulenceModel::None; }
    void setMainKernel(const std::string& k) { mainKernel = k; }
    const std::string& getMainKernel() const { return mainKernel; }
    /* Creates the level structs: omega = 1 / (3 nu 2^level + 1/2), forceFactor = 2^-level
     * (Parameter.cpp:395-418), parity even. */
    void initLBMSimulationParameter();
    std::shared_ptr<LBMSimulationParameter> getParD(int level) const { return parD.at(level); }
    bool getKernelNeedsFluidNodeIndicesToRun() const { return needsFluidNodeIndices; }
    void setKernelNeedsFluidNodeIndicesToRun(bool v) { needsFluidNodeIndices = v; }

private:
    int maxLevel = 0;
    double viscosity = 0.001;
    float limiters[3] = {0.01f, 0.01f, 0.01f};
    float forces[3] = {0.f, 0.f, 0.f};
    vf_storage storage = VF_STORAGE_FP32;
    TurbulenceModel turbulenceModel = TurbulenceModel::None;
    std::string mainKernel = "K17CompressibleNavierStokes";
    bool needsFluidNodeIndices = true;
    std::vector<std::shared_ptr<LBMSimulationParameter>> parD;
};

/* Device memory of one level, allocated from host arrays (the CudaMemoryManager role,
 * src/gpu/core/Cuda/CudaMemoryManager.cpp:169). Filling a LBMSimulationParameter's
 * pointers is left to attach(). */
class DeviceLevel {
public:
    struct Host {
        uint64_t numberOfNodes = 0;
        int nx = 0, ny = 0, nz = 0; /* dense periodic box extents (0 for a sparse grid) */
        std::vector<uint32_t> neighborX, neighborY, neighborZ, typeOfGridNode, fluidNodeIndices;
        std::vector<uint8_t> fluidMask; /* optional (dense box with solid nodes) */
    };
    struct HostBC {
        std::vector<int32_t> k, kN;
        std::vector<float> q27, Vx, Vy, Vz, RhoBC;
        int32_t direction = 0;
    };
    struct HostCells {
        std::vector<uint32_t> coarse, fine;
        std::vector<float> x, y, z; /* optional offsets */
    };

    explicit DeviceLevel(const Host& host, vf_storage storage);
    ~DeviceLevel();
    DeviceLevel(const DeviceLevel&) = delete;
    DeviceLevel& operator=(const DeviceLevel&) = delete;

    void attach(LBMSimulationParameter& p) const; /* points p at this level's memory */
    void setBC(vf_q_boundary_conditions& bc, const HostBC& host); /* plain or directional BC list */
    /* throws std::invalid_argument for a BC list setBC rejects: sizes that do not match, a node
     * outside [0, numberOfNodes) or a node listed twice (the BC kernels write each
     * (node, direction) cell from one thread) */
    static void checkBC(const HostBC& host, uint64_t numberOfNodes);
    /* cells of the interface from this (coarse) level to a finer level of fineNumberOfNodes */
    void setCells(InterpolationCells& cells, ICellNeigh& neigh, bool& hasOffsets, const HostCells& host,
                  uint64_t fineNumberOfNodes);
    /* drho, vx, vy, vz, pressure of every node (host copies, after a device sync) */
    std::vector<float> download(const float* devicePtr) const;

private:
    void* upload(const void* src, size_t bytes);
    std::vector<void*> owned;
    Host shape;
    vf_storage storage;
    void* distributions = nullptr;
    uint32_t *nbx = nullptr, *nby = nullptr, *nbz = nullptr, *geo = nullptr, *fluid = nullptr;
    uint8_t* mask = nullptr;
    float *rho = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr, *press = nullptr;

public:
    const float* rhoDevice() const { return rho; }
    const float* vxDevice() const { return vx; }
    const float* vyDevice() const { return vy; }
    const float* vzDevice() const { return vz; }
    const float* pressureDevice() const { return press; }
};

/* ---- checkpoint / restart (RestartObject, src/gpu/core/Restart/RestartObject.h:45-80) ------ */
/* The reference's restart file name: fname + "_Restart_" + myID + "_" + step (Simulation.cpp:93-96). */
std::string getFileName(const std::string& fname, int step, int myID);

/* Population buffers of every level, one float vector each (27 * numberOfNodes values in the
 * device slot layout; FP16/BF16 levels widened to float and narrowed back with round to
 * nearest even). serialize downloads and writes (NaN written as 0); deserialize reads and
 * uploads, after checking the level count and sizes (the reference copies unchecked). Neither
 * touches isEvenTimestep: the reference's PreProcessor leaves it even before a restart, so
 * checkpoints written after an even number of coarse steps resume exactly. */
class RestartObject {
public:
    virtual ~RestartObject() = default;
    void deserialize(const std::string& filename, Parameter& para);
    void serialize(const std::string& filename, const Parameter& para);
    void delete_restart_file(const std::string& filename); /* prints a warning when absent */
    std::vector<std::vector<float>> fs;
    virtual void serialize_internal(const std::string& filename) = 0;
    virtual void deserialize_internal(const std::string& filename) = 0;

protected:
    virtual std::string getFileExtension() const = 0;
};

/* ".txt": one line per level, values printed with 6 significant digits (lossy, as the reference) */
class ASCIIRestartObject : public RestartObject {
public:
    void serialize_internal(const std::string& filename) override;
    void deserialize_internal(const std::string& filename) override;

protected:
    std::string getFileExtension() const override;
};

/* ".bin": int32 level count, then per level int32 count and the floats (exact) */
class BinaryRestartObject : public RestartObject {
public:
    void serialize_internal(const std::string& filename) override;
    void deserialize_internal(const std::string& filename) override;

protected:
    std::string getFileExtension() const override;
};

/* ---- kernels ---------------------------------------------------------------------------- */
class Kernel {
public:
    virtual ~Kernel() = default;
    virtual void run() = 0;
    virtual void runOnIndices(const unsigned int* indices, unsigned int size_indices,
                              CollisionTemplate collisionTemplate, hipStream_t stream = nullptr) = 0;
    virtual std::vector<PreProcessorType> getPreProcessorTypes() = 0;
};

class K17CompressibleNavierStokes : public Kernel {
public:
    static std::shared_ptr<K17CompressibleNavierStokes> getNewInstance(std::shared_ptr<Parameter> para, int level);
    void run() override;
    void runOnIndices(const unsigned int* indices, unsigned int size_indices, CollisionTemplate collisionTemplate,
                      hipStream_t stream = nullptr) override;
    std::vector<PreProcessorType> getPreProcessorTypes() override
    {
        return {PreProcessorType::InitNavierStokesCompressible};
    }
    bool getKernelUsesFluidNodeIndices() const { return true; }

private:
    K17CompressibleNavierStokes(std::shared_ptr<Parameter> para, int level) : para(std::move(para)), level(level) {}
    std::shared_ptr<Parameter> para;
    int level;
};

namespace collisionKernel::compressible {
static const std::string K17CompressibleNavierStokes = "K17CompressibleNavierStokes";
}

class KernelFactory {
public:
    std::vector<std::shared_ptr<Kernel>> makeKernels(std::shared_ptr<Parameter> para);
    std::shared_ptr<Kernel> makeKernel(std::shared_ptr<Parameter> para, const std::string& kernel, int level);
};

/* InitNavierStokesCompressible: equilibrium from the level's rho/velocity arrays. */
class PreProcessor {
public:
    explicit PreProcessor(std::shared_ptr<Parameter> para) : para(std::move(para)) {}
    void init(int level);

private:
    std::shared_ptr<Parameter> para;
};

/* ---- boundary conditions ----------------------------------------------------------------- */
using boundaryCondition = std::function<void(LBMSimulationParameter*, QforBoundaryConditions*)>;
using boundaryConditionDirectional = std::function<void(LBMSimulationParameter*, QforDirectionalBoundaryCondition*)>;

class BoundaryConditionFactory {
public:
    enum class VelocityBC {
        VelocityBounceBack,
        VelocityInterpolatedIncompressible,
        VelocityInterpolatedCompressible,
        VelocityWithPressureInterpolatedCompressible,
        NotSpecified
    };
    enum class NoSlipBC {
        NoSlipDelayBounceBack,
        NoSlipBounceBack,
        NoSlipInterpolatedIncompressible,
        NoSlipInterpolatedCompressible,
    };
    enum class SlipBC { SlipCompressible, SlipTurbulentViscosityCompressible, NotSpecified };
    enum class PressureBC {
        PressureNonEquilibriumIncompressible,
        PressureNonEquilibriumCompressible,
        OutflowNonReflective,
        OutflowNonReflectivePressureCorrection,
        NotSpecified
    };

    void setVelocityBoundaryCondition(VelocityBC bc) { velocityBoundaryCondition = bc; }
    void setNoSlipBoundaryCondition(NoSlipBC bc) { noSlipBoundaryCondition = bc; }
    void setSlipBoundaryCondition(SlipBC bc) { slipBoundaryCondition = bc; }
    void setPressureBoundaryCondition(PressureBC bc) { pressureBoundaryCondition = bc; }
    void setGeometryBoundaryCondition(std::variant<VelocityBC, NoSlipBC, SlipBC> bc) { geometryBoundaryCondition = bc; }

    /* nullptr for NotSpecified; an empty function for NoSlipDelayBounceBack; throws
     * std::runtime_error for variants outside the K17 compressible path. */
    boundaryCondition getVelocityBoundaryConditionPost(bool isGeometryBC = false) const;
    boundaryCondition getNoSlipBoundaryConditionPost(bool isGeometryBC = false) const;
    boundaryCondition getSlipBoundaryConditionPost(bool isGeometryBC = false) const;
    boundaryCondition getGeometryBoundaryConditionPost() const;
    std::variant<boundaryCondition, boundaryConditionDirectional> getPressureBoundaryConditionPre() const;
    bool hasDirectionalPressureBoundaryCondition() const;

private:
    VelocityBC velocityBoundaryCondition = VelocityBC::NotSpecified;
    NoSlipBC noSlipBoundaryCondition = NoSlipBC::NoSlipDelayBounceBack;
    SlipBC slipBoundaryCondition = SlipBC::NotSpecified;
    PressureBC pressureBoundaryCondition = PressureBC::NotSpecified;
    std::variant<VelocityBC, NoSlipBC, SlipBC> geometryBoundaryCondition = NoSlipBC::NoSlipDelayBounceBack;
};

class BoundaryConditionKernelManager {
public:
    BoundaryConditionKernelManager(std::shared_ptr<Parameter> para, const BoundaryConditionFactory* bcFactory);
    void runVelocityBCKernelPost(int level) const;
    void runNoSlipBCKernelPost(int level) const;
    void runSlipBCKernelPost(int level) const;
    void runGeometryBCKernelPost(int level) const;
    void runPressureBCKernelPre(int level) const;

private:
    std::shared_ptr<Parameter> para;
    boundaryCondition velocityBoundaryConditionPost, noSlipBoundaryConditionPost, slipBoundaryConditionPost,
        geometryBoundaryConditionPost, pressureBoundaryConditionPre;
    boundaryConditionDirectional directionalPressureBoundaryConditionPre;
};

/* ---- grid refinement --------------------------------------------------------------------- */
using gridScaling =
    std::function<void(LBMSimulationParameter*, LBMSimulationParameter*, InterpolationCells*, ICellNeigh&, hipStream_t)>;

class GridScalingFactory {
public:
    enum class GridScaling { ScaleCompressible, NotSpecified };
    void setScalingFactory(GridScaling type) { gridScalingType = type; }
    gridScaling getGridScalingFC(bool hasTurbulentViscosity) const;
    gridScaling getGridScalingCF(bool hasTurbulentViscosity) const;

private:
    GridScaling gridScalingType = GridScaling::NotSpecified;
};

class GridScalingKernelManager {
public:
    GridScalingKernelManager(std::shared_ptr<Parameter> para, GridScalingFactory* gridScalingFactory);
    void runFineToCoarseKernelLB(int level, InterpolationCells* fineToCoarse, ICellNeigh& neighborFineToCoarse,
                                 hipStream_t stream) const;
    void runCoarseToFineKernelLB(int level, InterpolationCells* coarseToFine, ICellNeigh& neighborCoarseToFine,
                                 hipStream_t stream) const;

private:
    std::shared_ptr<Parameter> para;
    gridScaling scalingFineToCoarse = nullptr;
    gridScaling scalingCoarseToFine = nullptr;
};

/* ---- the time step ----------------------------------------------------------------------- */
class UpdateGrid27 {
public:
    UpdateGrid27(std::shared_ptr<Parameter> para, std::vector<std::shared_ptr<Kernel>> kernels,
                 const BoundaryConditionFactory* bcFactory, GridScalingFactory* scalingFactory);
    /* one step of `level` (and two of every finer level before it) */
    void updateGrid(int level, unsigned int t);
    /* `steps` coarse steps; with graphs (the default) pairs of steps replay one HIP graph
     * (no counterpart in the reference, whose loop launches every kernel from the host) */
    void run(unsigned int steps);
    void setUseGraphs(bool on) { useGraphs = on; }

private:
    void collision(int level);
    void postCollisionBC(int level);
    void swapBetweenEvenAndOddTimestep(int level);
    void preCollisionBC(int level);
    void refinement(int level);

    std::shared_ptr<Parameter> para;
    std::vector<std::shared_ptr<Kernel>> kernels;
    BoundaryConditionKernelManager bcKernelManager;
    GridScalingKernelManager gridScalingKernelManager;
    unsigned int timestep = 0;
    bool useGraphs = true;
};

} // namespace vfamd

#endif /* VIRTUALFLUIDS_AMD_HPP */
