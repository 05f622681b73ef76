// Checkpoint and restart files of the population buffers (host side of the C++ layer).
//
// Same files as the reference's RestartObject (src/gpu/core/Restart/RestartObject.cpp:43-201):
// one vector of 27 * numberOfNodes `real` (float) values per level, in the slot layout of the
// device buffer, NaN written as 0.
//   binary (.bin): int32 number of levels, then per level int32 count and count floats;
//   ASCII  (.txt): one line per level, every value followed by a space, printed with the
//                  stream defaults (6 significant digits).
// File name: fname + "_Restart_" + processID + "_" + step (Simulation.cpp:93-96).
// Levels stored in FP16 or BF16 are widened to float on the way out and rounded to nearest
// even on the way back, so a restart of a reduced-storage run is exact.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>

#include "virtualfluids_amd.hpp"

namespace vfamd {

namespace {

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

size_t storage_bytes(int32_t s)
{
    if (s != VF_STORAGE_FP32 && s != VF_STORAGE_FP16 && s != VF_STORAGE_BF16)
        throw std::runtime_error("restart: unknown storage type " + std::to_string(s));
    return s == VF_STORAGE_FP32 ? 4 : 2;
}

float widen(int32_t s, const unsigned char* p)
{
    if (s == VF_STORAGE_FP32) {
        float v;
        std::memcpy(&v, p, 4);
        return v;
    }
    uint16_t h;
    std::memcpy(&h, p, 2);
    if (s == VF_STORAGE_BF16) {
        const uint32_t w = (uint32_t)h << 16;
        float v;
        std::memcpy(&v, &w, 4);
        return v;
    }
    _Float16 v;
    std::memcpy(&v, &h, 2);
    return (float)v;
}

void narrow(int32_t s, float v, unsigned char* p)
{
    if (s == VF_STORAGE_FP32) {
        std::memcpy(p, &v, 4);
    } else if (s == VF_STORAGE_BF16) {
        const __bf16 b = (__bf16)v; // round to nearest even
        std::memcpy(p, &b, 2);
    } else {
        const _Float16 h = (_Float16)v; // round to nearest even
        std::memcpy(p, &h, 2);
    }
}

} // namespace

std::string getFileName(const std::string& fname, int step, int myID)
{
    return fname + "_Restart_" + std::to_string(myID) + "_" + std::to_string(step);
}

void RestartObject::serialize(const std::string& filename, const Parameter& para)
{
    fs.clear();
    for (int level = para.getCoarse(); level <= para.getFine(); ++level) {
        const auto p = para.getParD(level);
        const size_t n = 27 * (size_t)p->numberOfNodes, w = storage_bytes(p->storage);
        std::vector<unsigned char> raw(n * w);
        hip_check(hipDeviceSynchronize(), "restart: synchronize");
        hip_check(hipMemcpy(raw.data(), p->distributions, raw.size(), hipMemcpyDeviceToHost), "restart: download");
        std::vector<float> v(n);
        for (size_t i = 0; i < n; ++i) {
            const float x = widen(p->storage, raw.data() + i * w);
            v[i] = std::isnan(x) ? 0.0f : x;
        }
        fs.push_back(std::move(v));
    }
    serialize_internal(filename);
}

void RestartObject::deserialize(const std::string& filename, Parameter& para)
{
    fs.clear();
    deserialize_internal(filename);
    // unlike the reference, the sizes are checked before anything is copied
    if ((int)fs.size() < para.getFine() + 1)
        throw std::runtime_error("restart file " + filename + " holds " + std::to_string(fs.size()) +
                                 " level(s), the simulation has " + std::to_string(para.getFine() + 1));
    for (int level = para.getCoarse(); level <= para.getFine(); ++level) {
        const auto p = para.getParD(level);
        const size_t n = 27 * (size_t)p->numberOfNodes, w = storage_bytes(p->storage);
        if (fs[level].size() != n)
            throw std::runtime_error("restart file " + filename + ": level " + std::to_string(level) + " holds " +
                                     std::to_string(fs[level].size()) + " values, expected " + std::to_string(n));
        std::vector<unsigned char> raw(n * w);
        for (size_t i = 0; i < n; ++i) narrow(p->storage, fs[level][i], raw.data() + i * w);
        hip_check(hipMemcpy(p->distributions, raw.data(), raw.size(), hipMemcpyHostToDevice), "restart: upload");
    }
}

void RestartObject::delete_restart_file(const std::string& filename)
{
    const std::string file = filename + getFileExtension();
    std::error_code ec;
    if (!std::filesystem::remove(file, ec))
        std::cout << "[warning] Could not delete the restart object file \"" << file << "\"." << std::endl;
}

std::string ASCIIRestartObject::getFileExtension() const { return ".txt"; }

void ASCIIRestartObject::serialize_internal(const std::string& filename)
{
    std::ofstream stream(filename + getFileExtension());
    if (!stream) throw std::runtime_error("couldn't open file " + filename + getFileExtension());
    for (const auto& level : fs) {
        for (const float v : level) stream << v << " ";
        stream << "\n";
    }
}

void ASCIIRestartObject::deserialize_internal(const std::string& filename)
{
    std::ifstream stream(filename + getFileExtension());
    if (!stream.is_open()) throw std::runtime_error("couldn't open check point file " + filename);
    std::string line;
    while (std::getline(stream, line)) {
        std::vector<float> values;
        std::stringstream ls(line);
        float v;
        while (ls >> v) values.push_back(v);
        fs.push_back(std::move(values));
    }
}

std::string BinaryRestartObject::getFileExtension() const { return ".bin"; }

void BinaryRestartObject::serialize_internal(const std::string& filename)
{
    std::ofstream stream(filename + getFileExtension(), std::ios_base::binary);
    if (!stream) throw std::runtime_error("couldn't open file " + filename + getFileExtension());
    const int32_t levels = (int32_t)fs.size();
    stream.write(reinterpret_cast<const char*>(&levels), sizeof(levels));
    for (const auto& level : fs) {
        if (level.size() > (size_t)INT32_MAX) throw std::runtime_error("restart: a level exceeds 2^31 values");
        const int32_t count = (int32_t)level.size();
        stream.write(reinterpret_cast<const char*>(&count), sizeof(count));
        stream.write(reinterpret_cast<const char*>(level.data()), (std::streamsize)(level.size() * sizeof(float)));
    }
}

void BinaryRestartObject::deserialize_internal(const std::string& filename)
{
    std::ifstream stream(filename + getFileExtension(), std::ios_base::binary);
    if (!stream.is_open()) throw std::runtime_error("couldn't open check point file " + filename);
    int32_t levels = 0;
    stream.read(reinterpret_cast<char*>(&levels), sizeof(levels));
    if (!stream || levels < 0) throw std::runtime_error("restart: truncated file " + filename);
    fs.resize((size_t)levels);
    for (auto& level : fs) {
        int32_t count = 0;
        stream.read(reinterpret_cast<char*>(&count), sizeof(count));
        if (!stream || count < 0) throw std::runtime_error("restart: truncated file " + filename);
        level.resize((size_t)count);
        stream.read(reinterpret_cast<char*>(level.data()), (std::streamsize)(level.size() * sizeof(float)));
        if (!stream) throw std::runtime_error("restart: truncated file " + filename);
    }
}

} // namespace vfamd
