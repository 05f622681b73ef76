// Tests of the C++ host layer (include/virtualfluids_amd.hpp).
//
//   test_host factories        factory and manager semantics, no GPU needed; mirrors
//                              tests/unit-tests/gpu/core/BoundaryConditions/BoundaryConditionFactoryTest.cpp,
//                              BCKernelManagerTest.cpp and
//                              tests/unit-tests/gpu/core/GridScaling/GridScalingKernelManagerTest.cpp
//   test_host run DIR STEPS    runs a case written by tests/cpp_case.py through UpdateGrid27 on
//                              the GPU and writes the macroscopic fields of every level to DIR
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>

#include "virtualfluids_amd.hpp"

using namespace vfamd;

static int failures = 0;
#define EXPECT(cond)                                                                     \
    do {                                                                                 \
        if (!(cond)) {                                                                   \
            std::fprintf(stderr, "%s:%d: EXPECT(%s) failed\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                                  \
        }                                                                                \
    } while (0)
#define EXPECT_THROW(stmt, Ex)                                                          \
    do {                                                                                \
        bool thrown = false;                                                            \
        try {                                                                           \
            stmt;                                                                       \
        } catch (const Ex&) {                                                           \
            thrown = true;                                                              \
        }                                                                               \
        if (!thrown) {                                                                  \
            std::fprintf(stderr, "%s:%d: %s did not throw\n", __FILE__, __LINE__, #stmt); \
            ++failures;                                                                 \
        }                                                                               \
    } while (0)
#define EXPECT_NO_THROW(stmt)                                                                      \
    do {                                                                                           \
        try {                                                                                      \
            stmt;                                                                                  \
        } catch (const std::exception& e) {                                                        \
            std::fprintf(stderr, "%s:%d: %s threw: %s\n", __FILE__, __LINE__, #stmt, e.what());    \
            ++failures;                                                                            \
        }                                                                                          \
    } while (0)

static std::shared_ptr<Parameter> make_parameter(int maxLevel)
{
    auto para = std::make_shared<Parameter>();
    para->setMaxLevel(maxLevel);
    para->initLBMSimulationParameter();
    return para;
}

// ---- BoundaryConditionFactoryTest ---------------------------------------------------------
static void bc_factory_tests()
{
    using F = BoundaryConditionFactory;
    {
        F f;
        auto bc = f.getVelocityBoundaryConditionPost();
        EXPECT(!bc);
        EXPECT_THROW(bc(nullptr, nullptr), std::bad_function_call);
    }
    {
        F f;
        auto bc = f.getNoSlipBoundaryConditionPost();
        EXPECT(bc);
        EXPECT_NO_THROW(bc(nullptr, nullptr)); // empty function: NoSlipDelayBounceBack
    }
    {
        F f;
        auto bc = f.getSlipBoundaryConditionPost();
        EXPECT(!bc);
        EXPECT_THROW(bc(nullptr, nullptr), std::bad_function_call);
    }
    {
        F f;
        auto v = f.getPressureBoundaryConditionPre();
        EXPECT(std::holds_alternative<boundaryCondition>(v));
        EXPECT(!std::get<boundaryCondition>(v));
        EXPECT(!f.hasDirectionalPressureBoundaryCondition());
    }
    {
        F f;
        auto bc = f.getGeometryBoundaryConditionPost();
        EXPECT(bc);
        EXPECT_NO_THROW(bc(nullptr, nullptr));
    }
    {
        F f;
        for (auto v : {F::VelocityBC::VelocityBounceBack, F::VelocityBC::VelocityInterpolatedCompressible,
                       F::VelocityBC::VelocityWithPressureInterpolatedCompressible}) {
            f.setVelocityBoundaryCondition(v);
            EXPECT(f.getVelocityBoundaryConditionPost());
            f.setGeometryBoundaryCondition(v);
            EXPECT(f.getGeometryBoundaryConditionPost());
        }
        f.setVelocityBoundaryCondition(F::VelocityBC::VelocityInterpolatedIncompressible);
        EXPECT_THROW(f.getVelocityBoundaryConditionPost(), std::runtime_error);
        for (auto v : {F::NoSlipBC::NoSlipBounceBack, F::NoSlipBC::NoSlipInterpolatedCompressible}) {
            f.setNoSlipBoundaryCondition(v);
            EXPECT(f.getNoSlipBoundaryConditionPost());
        }
        f.setNoSlipBoundaryCondition(F::NoSlipBC::NoSlipInterpolatedIncompressible);
        EXPECT_THROW(f.getNoSlipBoundaryConditionPost(), std::runtime_error);
        f.setPressureBoundaryCondition(F::PressureBC::PressureNonEquilibriumCompressible);
        EXPECT(f.hasDirectionalPressureBoundaryCondition());
        auto v = f.getPressureBoundaryConditionPre();
        EXPECT(std::holds_alternative<boundaryConditionDirectional>(v));
        EXPECT(std::get<boundaryConditionDirectional>(v));
        f.setPressureBoundaryCondition(F::PressureBC::OutflowNonReflective);
        EXPECT(f.hasDirectionalPressureBoundaryCondition());
        EXPECT_THROW(f.getPressureBoundaryConditionPre(), std::runtime_error);
    }
}

// ---- BCKernelManagerTest (BCsNotSpecified) ------------------------------------------------
static void bc_manager_tests()
{
    BoundaryConditionFactory f;
    {
        auto para = make_parameter(0);
        EXPECT_NO_THROW(BoundaryConditionKernelManager(para, &f));
    }
    {
        auto para = make_parameter(0);
        para->getParD(0)->velocityBC.numberOfBCnodes = 1;
        EXPECT_THROW(BoundaryConditionKernelManager(para, &f), std::runtime_error);
    }
    {
        auto para = make_parameter(0);
        para->getParD(0)->noSlipBC.numberOfBCnodes = 1; // a default is specified
        EXPECT_NO_THROW(BoundaryConditionKernelManager(para, &f));
    }
    {
        auto para = make_parameter(0);
        para->getParD(0)->slipBC.numberOfBCnodes = 1;
        EXPECT_THROW(BoundaryConditionKernelManager(para, &f), std::runtime_error);
    }
    {
        auto para = make_parameter(0);
        para->getParD(0)->pressureBC.numberOfBCnodes = 1;
        EXPECT_THROW(BoundaryConditionKernelManager(para, &f), std::runtime_error);
    }
    {
        auto para = make_parameter(0);
        para->getParD(0)->pressureBCDirectional.emplace_back();
        EXPECT_NO_THROW(BoundaryConditionKernelManager(para, &f)); // an empty directional list
        para->getParD(0)->pressureBCDirectional.back().numberOfBCnodes = 1;
        EXPECT_THROW(BoundaryConditionKernelManager(para, &f), std::runtime_error);
    }
    {
        auto para = make_parameter(0);
        para->getParD(0)->geometryBC.numberOfBCnodes = 1; // a default is specified
        EXPECT_NO_THROW(BoundaryConditionKernelManager(para, &f));
    }
    {
        BoundaryConditionFactory g;
        g.setVelocityBoundaryCondition(BoundaryConditionFactory::VelocityBC::VelocityInterpolatedCompressible);
        g.setPressureBoundaryCondition(BoundaryConditionFactory::PressureBC::PressureNonEquilibriumCompressible);
        auto para = make_parameter(0);
        para->getParD(0)->velocityBC.numberOfBCnodes = 1;
        para->getParD(0)->pressureBCDirectional.emplace_back();
        para->getParD(0)->pressureBCDirectional.back().numberOfBCnodes = 1;
        EXPECT_NO_THROW(BoundaryConditionKernelManager(para, &g));
        // runXXX with no nodes calls nothing (and so needs no GPU)
        para->getParD(0)->velocityBC.numberOfBCnodes = 0;
        para->getParD(0)->pressureBCDirectional.back().numberOfBCnodes = 0;
        BoundaryConditionKernelManager m(para, &g);
        EXPECT_NO_THROW(m.runVelocityBCKernelPost(0));
        EXPECT_NO_THROW(m.runPressureBCKernelPre(0));
    }
}

// ---- BC lists checked on the host before any upload -------------------------------------------
static void bc_list_tests()
{
    DeviceLevel::HostBC bc;
    bc.k = {3, 7, 5};
    bc.q27.assign(27 * 3, -1.0f);
    EXPECT_NO_THROW(DeviceLevel::checkBC(bc, 8));
    EXPECT_THROW(DeviceLevel::checkBC(bc, 7), std::invalid_argument); // node 7 out of range
    bc.k = {3, 7, 3};
    EXPECT_THROW(DeviceLevel::checkBC(bc, 8), std::invalid_argument); // node 3 twice
    bc.k = {3, 7, 5};
    bc.Vx = {0.1f, 0.2f};
    EXPECT_THROW(DeviceLevel::checkBC(bc, 8), std::invalid_argument); // one value short
    bc.Vx.clear();
    bc.q27.pop_back();
    EXPECT_THROW(DeviceLevel::checkBC(bc, 8), std::invalid_argument);
}

// ---- GridScalingKernelManagerTest_Initialization -------------------------------------------
static void scaling_manager_tests()
{
    GridScalingFactory f;
    EXPECT_NO_THROW(GridScalingKernelManager(make_parameter(0), nullptr));
    EXPECT_NO_THROW(GridScalingKernelManager(make_parameter(0), &f));
    {
        GridScalingFactory g;
        g.setScalingFactory(GridScalingFactory::GridScaling::ScaleCompressible);
        EXPECT_NO_THROW(GridScalingKernelManager(make_parameter(0), &g));
        EXPECT_NO_THROW(GridScalingKernelManager(make_parameter(2), &g));
        EXPECT(g.getGridScalingFC(false));
        EXPECT(g.getGridScalingCF(false));
        EXPECT_THROW(g.getGridScalingFC(true), std::runtime_error);
    }
    EXPECT_THROW(GridScalingKernelManager(make_parameter(2), nullptr), std::runtime_error);
    EXPECT_NO_THROW(GridScalingKernelManager(make_parameter(2), &f));
    {
        auto para = make_parameter(2);
        para->getParD(0)->fineToCoarse.numberOfCells = 100;
        EXPECT_THROW(GridScalingKernelManager(para, &f), std::runtime_error);
    }
    {
        auto para = make_parameter(2);
        para->getParD(0)->coarseToFine.numberOfCells = 100;
        EXPECT_THROW(GridScalingKernelManager(para, &f), std::runtime_error);
    }
}

// ---- KernelFactory and Parameter -------------------------------------------------------------
static void kernel_factory_tests()
{
    KernelFactory kf;
    auto para = make_parameter(1);
    auto kernels = kf.makeKernels(para);
    EXPECT(kernels.size() == 2);
    EXPECT(kernels[0]->getPreProcessorTypes().size() == 1);
    EXPECT(para->getKernelNeedsFluidNodeIndicesToRun());
    EXPECT_THROW(kf.makeKernel(para, "NoSuchKernel", 0), std::runtime_error);
    para->setTurbulenceModel(TurbulenceModel::Smagorinsky);
    EXPECT_THROW(kf.makeKernel(para, collisionKernel::compressible::K17CompressibleNavierStokes, 0),
                 std::runtime_error);
    // omega per level: 1 / (3 nu 2^level + 1/2), forceFactor 2^-level
    auto p = std::make_shared<Parameter>();
    p->setMaxLevel(2);
    p->setViscosityLB(0.01);
    p->initLBMSimulationParameter();
    EXPECT(std::fabs(p->getParD(0)->omega - 1.0f / 0.53f) < 1e-6f);
    EXPECT(std::fabs(p->getParD(2)->omega - 1.0f / 0.62f) < 1e-6f);
    EXPECT(p->getParD(2)->forceFactor == 0.25f);
    EXPECT(p->getParD(1)->isEvenTimestep == 1);
    // UpdateGrid27 hands the levels of a hierarchy the sum of their population bytes (the
    // collision picks its kernel from it); a level the caller set keeps its value
    {
        auto hp = make_parameter(1);
        hp->getParD(0)->numberOfNodes = 1000;
        hp->getParD(0)->storage = VF_STORAGE_FP32;
        hp->getParD(1)->numberOfNodes = 500;
        hp->getParD(1)->storage = VF_STORAGE_FP16;
        hp->getParD(1)->workingSetBytes = 7;
        BoundaryConditionFactory bf;
        GridScalingFactory gf;
        UpdateGrid27 ug(hp, kf.makeKernels(hp), &bf, &gf);
        EXPECT(hp->getParD(0)->workingSetBytes == 27ull * (1000 * 4 + 500 * 2));
        EXPECT(hp->getParD(1)->workingSetBytes == 7);
        auto single = make_parameter(0);
        single->getParD(0)->numberOfNodes = 1000;
        UpdateGrid27 us(single, kf.makeKernels(single), &bf, &gf);
        EXPECT(single->getParD(0)->workingSetBytes == 0);
    }
    // the C ABI rejects a null level before touching the GPU
    EXPECT(vf_k17_run(nullptr, nullptr) == VF_ERR_INVALID_ARGUMENT);
    EXPECT_THROW(check(vf_k17_run(nullptr, nullptr)), std::runtime_error);
}

// ---- run: a case directory written by tests/cpp_case.py ------------------------------------
template <class T> static std::vector<T> read_array(const std::string& path, bool required = true)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        if (required) throw std::runtime_error("missing " + path);
        return {};
    }
    in.seekg(0, std::ios::end);
    const size_t bytes = (size_t)in.tellg();
    in.seekg(0);
    std::vector<T> v(bytes / sizeof(T));
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)bytes);
    return v;
}

template <class T> static void write_array(const std::string& path, const std::vector<T>& v)
{
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

static std::map<std::string, std::string> read_config(const std::string& path)
{
    std::map<std::string, std::string> cfg;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        const auto eq = line.find('=');
        if (eq != std::string::npos) cfg[line.substr(0, eq)] = line.substr(eq + 1);
    }
    return cfg;
}

// checkpoint / restart actions of a run: after `checkpointAt` steps write `file`, or start
// from `file` instead of the initial fields
struct RestartPlan {
    enum Kind { None, Checkpoint, Restart } kind = None;
    unsigned checkpointAt = 0;
    std::string file, format = "bin";
};

static std::unique_ptr<RestartObject> make_restart(const std::string& format)
{
    if (format == "txt") return std::make_unique<ASCIIRestartObject>();
    return std::make_unique<BinaryRestartObject>();
}

static int run_case(const std::string& dir, unsigned steps, const RestartPlan& plan = {})
{
    auto cfg = read_config(dir + "/config.txt");
    const int levels = std::stoi(cfg.at("levels"));
    auto para = std::make_shared<Parameter>();
    para->setMaxLevel(levels - 1);
    para->setViscosityLB(std::stod(cfg.at("viscosity")));
    const float lim = std::stof(cfg.at("limiter"));
    para->setQuadricLimiters(lim, lim, lim);
    para->setStorage((vf_storage)std::stoi(cfg.at("storage")));
    para->initLBMSimulationParameter();

    BoundaryConditionFactory bcf;
    using F = BoundaryConditionFactory;
    const std::map<std::string, F::VelocityBC> vel = {
        {"VelocityBounceBack", F::VelocityBC::VelocityBounceBack},
        {"VelocityInterpolatedCompressible", F::VelocityBC::VelocityInterpolatedCompressible},
        {"VelocityWithPressureInterpolatedCompressible", F::VelocityBC::VelocityWithPressureInterpolatedCompressible}};
    const std::map<std::string, F::NoSlipBC> ns = {
        {"NoSlipDelayBounceBack", F::NoSlipBC::NoSlipDelayBounceBack},
        {"NoSlipBounceBack", F::NoSlipBC::NoSlipBounceBack},
        {"NoSlipInterpolatedCompressible", F::NoSlipBC::NoSlipInterpolatedCompressible}};
    if (cfg.count("velocityBC")) bcf.setVelocityBoundaryCondition(vel.at(cfg["velocityBC"]));
    if (cfg.count("noSlipBC")) bcf.setNoSlipBoundaryCondition(ns.at(cfg["noSlipBC"]));
    if (cfg.count("pressureBC")) bcf.setPressureBoundaryCondition(F::PressureBC::PressureNonEquilibriumCompressible);
    GridScalingFactory gsf;
    gsf.setScalingFactory(GridScalingFactory::GridScaling::ScaleCompressible);

    std::vector<std::unique_ptr<DeviceLevel>> mem;
    for (int L = 0; L < levels; ++L) {
        const std::string pre = dir + "/L" + std::to_string(L) + "_";
        DeviceLevel::Host h;
        h.neighborX = read_array<uint32_t>(pre + "nbx.u32");
        h.neighborY = read_array<uint32_t>(pre + "nby.u32");
        h.neighborZ = read_array<uint32_t>(pre + "nbz.u32");
        h.typeOfGridNode = read_array<uint32_t>(pre + "geo.u32");
        h.fluidNodeIndices = read_array<uint32_t>(pre + "fluid.u32");
        h.fluidMask = read_array<uint8_t>(pre + "mask.u8", false);
        h.numberOfNodes = h.neighborX.size();
        h.nx = std::stoi(cfg.at("L" + std::to_string(L) + "_nx"));
        h.ny = std::stoi(cfg.at("L" + std::to_string(L) + "_ny"));
        h.nz = std::stoi(cfg.at("L" + std::to_string(L) + "_nz"));
        mem.push_back(std::make_unique<DeviceLevel>(h, (vf_storage)std::stoi(cfg.at("storage"))));
        auto p = para->getParD(L);
        mem.back()->attach(*p);
        for (const char* name : {"velocity", "noslip", "pressure"}) {
            DeviceLevel::HostBC bc;
            bc.k = read_array<int32_t>(pre + name + "_k.i32", false);
            if (bc.k.empty()) continue;
            bc.q27 = read_array<float>(pre + name + "_q27.f32");
            bc.Vx = read_array<float>(pre + name + "_vx.f32", false);
            bc.Vy = read_array<float>(pre + name + "_vy.f32", false);
            bc.Vz = read_array<float>(pre + name + "_vz.f32", false);
            bc.kN = read_array<int32_t>(pre + name + "_kN.i32", false);
            bc.RhoBC = read_array<float>(pre + name + "_rho.f32", false);
            if (!std::strcmp(name, "pressure")) {
                bc.direction = std::stoi(cfg.at("L" + std::to_string(L) + "_pressure_direction"));
                p->pressureBCDirectional.emplace_back();
                mem.back()->setBC(p->pressureBCDirectional.back(), bc);
            } else {
                mem.back()->setBC(!std::strcmp(name, "velocity") ? p->velocityBC : p->noSlipBC, bc);
            }
        }
        // initial macroscopic fields -> InitNavierStokesCompressible
        auto up = [&](const char* name, float* dst) {
            auto v = read_array<float>(pre + name);
            if (v.size() != h.numberOfNodes) throw std::runtime_error(pre + name + ": one value per node expected");
            check(hipMemcpy(dst, v.data(), v.size() * 4, hipMemcpyHostToDevice) == hipSuccess ? VF_OK : VF_ERR_HIP);
        };
        up("init_rho.f32", p->rho);
        up("init_vx.f32", p->velocityX);
        up("init_vy.f32", p->velocityY);
        up("init_vz.f32", p->velocityZ);
        PreProcessor(para).init(L);
    }
    for (int L = 0; L + 1 < levels; ++L) {
        const std::string pre = dir + "/L" + std::to_string(L) + "_";
        auto p = para->getParD(L);
        const uint64_t fineN = para->getParD(L + 1)->numberOfNodes;
        DeviceLevel::HostCells cf, fc;
        cf.coarse = read_array<uint32_t>(pre + "cf_coarse.u32");
        cf.fine = read_array<uint32_t>(pre + "cf_fine.u32");
        fc.coarse = read_array<uint32_t>(pre + "fc_coarse.u32");
        fc.fine = read_array<uint32_t>(pre + "fc_fine.u32");
        mem[L]->setCells(p->coarseToFine, p->neighborCoarseToFine, p->hasOffsetsCoarseToFine, cf, fineN);
        mem[L]->setCells(p->fineToCoarse, p->neighborFineToCoarse, p->hasOffsetsFineToCoarse, fc, fineN);
    }

    if (plan.kind == RestartPlan::Restart) make_restart(plan.format)->deserialize(plan.file, *para);
    KernelFactory kf;
    UpdateGrid27 update(para, kf.makeKernels(para), &bcf, &gsf);
    if (plan.kind == RestartPlan::Checkpoint) {
        update.run(plan.checkpointAt);
        make_restart(plan.format)->serialize(plan.file, *para);
        update.run(steps - plan.checkpointAt);
    } else {
        update.run(steps);
    }
    for (int L = 0; L < levels; ++L) {
        auto p = para->getParD(L);
        check(vf_CalculateMacroscopicQuantitiesCompressible(p.get(), p->stream));
        const std::string pre = dir + "/out_L" + std::to_string(L) + "_";
        write_array(pre + "rho.f32", mem[L]->download(mem[L]->rhoDevice()));
        write_array(pre + "vx.f32", mem[L]->download(mem[L]->vxDevice()));
        write_array(pre + "vy.f32", mem[L]->download(mem[L]->vyDevice()));
        write_array(pre + "vz.f32", mem[L]->download(mem[L]->vzDevice()));
        std::ofstream(pre + "parity.txt") << p->isEvenTimestep << "\n";
    }
    std::printf("ran %u coarse steps on %d level(s)\n", steps, levels);
    return 0;
}

// RestartObjectTests (tests/unit-tests/gpu/core/Restart/RestartObjectTests.cpp:39-88) restated
template <class T> static void restart_save_and_load(const std::string& dir)
{
    T write;
    write.fs = {{1, 2, 3}, {4, 5, 6}};
    const std::string name = dir + "/test_do_check_point";
    write.serialize_internal(name);
    T read;
    read.deserialize_internal(name);
    EXPECT(read.fs == write.fs);
    write.delete_restart_file(name);
}

static int restart_tests(const std::string& dir)
{
    restart_save_and_load<ASCIIRestartObject>(dir);
    restart_save_and_load<BinaryRestartObject>(dir);
    EXPECT(!std::ifstream(dir + "/test_do_check_point.bin") && !std::ifstream(dir + "/test_do_check_point.txt"));
    BinaryRestartObject b;
    b.delete_restart_file(dir + "/does_not_exist"); // prints a warning
    ASCIIRestartObject a;
    a.delete_restart_file(dir + "/does_not_exist");
    EXPECT(getFileName("out/cavity", 500, 0) == "out/cavity_Restart_0_500");
    // files for the comparison with the Python writer: two levels of values over many magnitudes
    std::vector<std::vector<float>> fs(2);
    for (int i = 0; i < 27 * 7; ++i) fs[0].push_back(0.0123456789f * std::sin(0.37f * (float)i) * std::pow(10.0f, (float)(i % 9) - 6.0f));
    for (int i = 0; i < 27 * 3; ++i) fs[1].push_back(-1.0f / (float)(i + 1));
    fs[1][5] = 0.0f;
    fs[1][6] = -0.0f;
    b.fs = fs;
    b.serialize_internal(dir + "/cpp");
    a.fs = fs;
    a.serialize_internal(dir + "/cpp");
    std::printf("restart: %d failure(s)\n", failures);
    return failures ? 1 : 0;
}

// the C structs' field offsets, for the ctypes binding's layout check
static int print_layout()
{
#define VF_FIELD(S, F) std::printf("%s.%s %zu\n", #S, #F, offsetof(S, F))
    VF_FIELD(vf_lbm_parameter, distributions);
    VF_FIELD(vf_lbm_parameter, storage);
    VF_FIELD(vf_lbm_parameter, isEvenTimestep);
    VF_FIELD(vf_lbm_parameter, numberOfNodes);
    VF_FIELD(vf_lbm_parameter, neighborX);
    VF_FIELD(vf_lbm_parameter, typeOfGridNode);
    VF_FIELD(vf_lbm_parameter, nx);
    VF_FIELD(vf_lbm_parameter, fluidMask);
    VF_FIELD(vf_lbm_parameter, omega);
    VF_FIELD(vf_lbm_parameter, quadricLimiters);
    VF_FIELD(vf_lbm_parameter, forces);
    VF_FIELD(vf_lbm_parameter, forceFactor);
    VF_FIELD(vf_lbm_parameter, bodyForceX);
    VF_FIELD(vf_lbm_parameter, rho);
    VF_FIELD(vf_lbm_parameter, pressure);
    VF_FIELD(vf_lbm_parameter, fluidNodeIndices);
    VF_FIELD(vf_lbm_parameter, numberOfFluidNodes);
    VF_FIELD(vf_lbm_parameter, porosity);
    VF_FIELD(vf_lbm_parameter, workingSetBytes);
    VF_FIELD(vf_lbm_parameter, kernelChoice);
    VF_FIELD(vf_q_boundary_conditions, k);
    VF_FIELD(vf_q_boundary_conditions, q27);
    VF_FIELD(vf_q_boundary_conditions, numberOfBCnodes);
    VF_FIELD(vf_q_boundary_conditions, Vx);
    VF_FIELD(vf_q_boundary_conditions, RhoBC);
    VF_FIELD(vf_q_boundary_conditions, direction);
    VF_FIELD(vf_interpolation_cells, numberOfCells);
    VF_FIELD(vf_interpolation_cell_neighbor, z);
#undef VF_FIELD
    std::printf("sizeof.vf_lbm_parameter %zu\n", sizeof(vf_lbm_parameter));
    std::printf("sizeof.vf_q_boundary_conditions %zu\n", sizeof(vf_q_boundary_conditions));
    return 0;
}

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "factories";
    if (mode == "layout") return print_layout();
    if (mode == "factories") {
        bc_factory_tests();
        bc_manager_tests();
        bc_list_tests();
        scaling_manager_tests();
        kernel_factory_tests();
        std::printf("%s: %d failure(s)\n", mode.c_str(), failures);
        return failures ? 1 : 0;
    }
    if (mode == "restart" && argc == 3) return restart_tests(argv[2]);
    if (mode == "run" && (argc == 4 || argc == 7 || argc == 8)) {
        try {
            RestartPlan plan;
            if (argc > 4) {
                const std::string action = argv[4];
                if (action == "checkpoint" && argc == 8) {
                    plan.kind = RestartPlan::Checkpoint;
                    plan.checkpointAt = (unsigned)std::stoul(argv[5]);
                    plan.file = argv[6];
                    plan.format = argv[7];
                } else if (action == "restart" && argc == 7) {
                    plan.kind = RestartPlan::Restart;
                    plan.file = argv[5];
                    plan.format = argv[6];
                } else {
                    throw std::invalid_argument("unknown restart action " + action);
                }
            }
            return run_case(argv[2], (unsigned)std::stoul(argv[3]), plan);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "run failed: %s\n", e.what());
            return 2;
        }
    }
    std::fprintf(stderr, "usage: test_host factories | layout | restart DIR | run DIR STEPS "
                         "[checkpoint T FILE bin|txt | restart FILE bin|txt]\n");
    return 64;
}
