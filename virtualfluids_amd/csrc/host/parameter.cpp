// Parameter levels and device memory of a level (host side of the C++ layer).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "virtualfluids_amd.hpp"

namespace vfamd {

void check(int status)
{
    if (status != VF_OK) {
        const char* msg = vf_get_last_error();
        throw std::runtime_error(msg && *msg ? msg : "virtualfluids_amd entry point failed");
    }
}

static void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void Parameter::setMaxLevel(int level)
{
    if (level < 0) throw std::invalid_argument("maxLevel must be >= 0");
    maxLevel = level;
}

void Parameter::setQuadricLimiters(float p, float m, float d)
{
    limiters[0] = p;
    limiters[1] = m;
    limiters[2] = d;
}

void Parameter::setForcesLB(float fx, float fy, float fz)
{
    forces[0] = fx;
    forces[1] = fy;
    forces[2] = fz;
}

// Parameter::initLBMSimulationParameter (src/gpu/core/Parameter/Parameter.cpp:395-418):
// viscosity 2^level nu, omega = 1 / (3 nu_level + 1/2) per level.
void Parameter::initLBMSimulationParameter()
{
    parD.clear();
    for (int level = 0; level <= maxLevel; ++level) {
        auto p = std::make_shared<LBMSimulationParameter>();
        p->viscosity = viscosity * std::pow(2.0, level);
        p->omega = float(1.0 / (3.0 * p->viscosity + 0.5));
        p->storage = storage;
        p->isEvenTimestep = 1;
        for (int i = 0; i < 3; ++i) {
            p->quadricLimiters[i] = limiters[i];
            p->forces[i] = forces[i];
        }
        p->forceFactor = float(std::pow(2.0, -level));
        parD.push_back(p);
    }
}

// ---- DeviceLevel --------------------------------------------------------------------------
void* DeviceLevel::upload(const void* src, size_t bytes)
{
    if (bytes == 0) return nullptr;
    void* d = nullptr;
    hip_check(hipMalloc(&d, bytes), "hipMalloc");
    owned.push_back(d);
    if (src) hip_check(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    else hip_check(hipMemset(d, 0, bytes), "hipMemset");
    return d;
}

static size_t storage_bytes(vf_storage s) { return s == VF_STORAGE_FP32 ? 4 : 2; }

DeviceLevel::DeviceLevel(const Host& host, vf_storage storage) : shape(host), storage(storage)
{
    const uint64_t N = host.numberOfNodes;
    if (N == 0 || host.neighborX.size() != N || host.neighborY.size() != N || host.neighborZ.size() != N)
        throw std::invalid_argument("DeviceLevel: neighbour arrays must hold numberOfNodes entries");
    if (!host.typeOfGridNode.empty() && host.typeOfGridNode.size() != N)
        throw std::invalid_argument("DeviceLevel: typeOfGridNode must hold numberOfNodes entries");
    if (!host.fluidMask.empty() && host.fluidMask.size() != N)
        throw std::invalid_argument("DeviceLevel: fluidMask must hold numberOfNodes entries");
    for (uint32_t k : host.fluidNodeIndices)
        if (k >= N) throw std::invalid_argument("DeviceLevel: fluid node index out of range");
    for (const auto* a : {&host.neighborX, &host.neighborY, &host.neighborZ})
        for (uint32_t k : *a)
            if (k >= N) throw std::invalid_argument("DeviceLevel: neighbour index out of range");
    distributions = upload(nullptr, 27 * N * storage_bytes(storage));
    nbx = (uint32_t*)upload(host.neighborX.data(), N * 4);
    nby = (uint32_t*)upload(host.neighborY.data(), N * 4);
    nbz = (uint32_t*)upload(host.neighborZ.data(), N * 4);
    if (!host.typeOfGridNode.empty()) geo = (uint32_t*)upload(host.typeOfGridNode.data(), N * 4);
    fluid = (uint32_t*)upload(host.fluidNodeIndices.data(), host.fluidNodeIndices.size() * 4);
    if (!host.fluidMask.empty()) mask = (uint8_t*)upload(host.fluidMask.data(), N);
    rho = (float*)upload(nullptr, N * 4);
    vx = (float*)upload(nullptr, N * 4);
    vy = (float*)upload(nullptr, N * 4);
    vz = (float*)upload(nullptr, N * 4);
    press = (float*)upload(nullptr, N * 4);
    shape.neighborX.clear();
    shape.neighborY.clear();
    shape.neighborZ.clear();
    shape.typeOfGridNode.clear();
    shape.fluidMask.clear();
}

DeviceLevel::~DeviceLevel()
{
    for (void* p : owned) (void)hipFree(p);
}

void DeviceLevel::attach(LBMSimulationParameter& p) const
{
    p.distributions = distributions;
    p.storage = storage;
    p.numberOfNodes = shape.numberOfNodes;
    p.neighborX = nbx;
    p.neighborY = nby;
    p.neighborZ = nbz;
    p.typeOfGridNode = geo;
    p.nx = shape.nx;
    p.ny = shape.ny;
    p.nz = shape.nz;
    p.fluidMask = mask;
    p.rho = rho;
    p.velocityX = vx;
    p.velocityY = vy;
    p.velocityZ = vz;
    p.pressure = press;
    p.fluidNodeIndices = fluid;
    p.numberOfFluidNodes = (uint32_t)shape.fluidNodeIndices.size();
}

void DeviceLevel::checkBC(const HostBC& host, uint64_t numberOfNodes)
{
    const size_t n = host.k.size();
    if (host.q27.size() != 27 * n) throw std::invalid_argument("setBC: q27 must hold 27 entries per BC node");
    if (!host.kN.empty() && host.kN.size() != n) throw std::invalid_argument("setBC: kN must hold one node per BC node");
    for (int32_t k : host.k)
        if (k < 0 || (uint64_t)k >= numberOfNodes) throw std::invalid_argument("setBC: BC node out of range");
    for (int32_t k : host.kN)
        if (k < 0 || (uint64_t)k >= numberOfNodes) throw std::invalid_argument("setBC: kN node out of range");
    for (const std::vector<float>* v : {&host.Vx, &host.Vy, &host.Vz, &host.RhoBC})
        if (!v->empty() && v->size() != n)
            throw std::invalid_argument("setBC: Vx, Vy, Vz and RhoBC must hold one value per BC node");
    // the BC kernels write each (node, direction) cell from one thread: a node listed twice would
    // be written by two threads in the same launch
    std::vector<int32_t> sorted(host.k);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        throw std::invalid_argument("setBC: a BC node is listed twice");
}

void DeviceLevel::setBC(vf_q_boundary_conditions& bc, const HostBC& host)
{
    checkBC(host, shape.numberOfNodes);
    const size_t n = host.k.size();
    auto opt = [&](const std::vector<float>& v) -> const float* {
        return v.empty() ? nullptr : (const float*)upload(v.data(), n * 4);
    };
    bc.k = (const int32_t*)upload(host.k.data(), n * 4);
    bc.kN = host.kN.empty() ? nullptr : (const int32_t*)upload(host.kN.data(), n * 4);
    bc.q27 = (const float*)upload(host.q27.data(), 27 * n * 4);
    bc.numberOfBCnodes = (uint32_t)n;
    bc.Vx = opt(host.Vx);
    bc.Vy = opt(host.Vy);
    bc.Vz = opt(host.Vz);
    bc.RhoBC = opt(host.RhoBC);
    bc.direction = host.direction;
    // layout hint of the adjoint kernels: consecutive nodes mostly neighbours in x
    size_t along = 0;
    for (size_t i = 1; i < n; ++i) along += host.k[i] == host.k[i - 1] + 1;
    bc.nodesAlongRows = n > 1 && 2 * along >= n - 1;
}

void DeviceLevel::setCells(InterpolationCells& cells, ICellNeigh& neigh, bool& hasOffsets, const HostCells& host,
                           uint64_t fineNumberOfNodes)
{
    const size_t n = host.coarse.size();
    if (host.fine.size() != n) throw std::invalid_argument("setCells: coarse and fine cell lists differ in length");
    for (uint32_t k : host.coarse)
        if (k >= shape.numberOfNodes) throw std::invalid_argument("setCells: coarse cell index out of range");
    for (uint32_t k : host.fine)
        if (k >= fineNumberOfNodes) throw std::invalid_argument("setCells: fine cell index out of range");
    cells.coarseCellIndices = (const uint32_t*)upload(host.coarse.data(), n * 4);
    cells.fineCellIndices = (const uint32_t*)upload(host.fine.data(), n * 4);
    cells.numberOfCells = (uint32_t)n;
    hasOffsets = !host.x.empty();
    if (hasOffsets) {
        if (host.x.size() != n || host.y.size() != n || host.z.size() != n)
            throw std::invalid_argument("setCells: offsets must hold one value per cell");
        neigh.x = (const float*)upload(host.x.data(), n * 4);
        neigh.y = (const float*)upload(host.y.data(), n * 4);
        neigh.z = (const float*)upload(host.z.data(), n * 4);
    }
}

std::vector<float> DeviceLevel::download(const float* devicePtr) const
{
    std::vector<float> out(shape.numberOfNodes);
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hip_check(hipMemcpy(out.data(), devicePtr, out.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    return out;
}

} // namespace vfamd
