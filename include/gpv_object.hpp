// gpv_object.hpp -- C++ facade with GPView's Object-level voxelization entry
// points (includes/Object.h:106-136), implemented on libgpv's C ABI.
//
//   reference                                     here
//   Object::ReadObject(char*)       :395-584      Object::ReadObject
//   Object::ReadOFFObject(char*)    :171-317      Object::ReadOFFObject
//   Object::CreateFlatTriangleData  :3496-3527    Object::CreateFlatTriangleData
//   Object::PerformVoxelization     :3077-3430    Object::PerformVoxelization
//   Object::SaveVoxelization        :2934-3075    Object::SaveVoxelization
//
// Error behaviour follows the reference: an unreadable input or output file
// prints to stderr and aborts (src/Object.cpp:409-413, 2988-2992).  A device
// error, which the reference only prints (src/CUDAUtilities.cpp:33-49),
// throws std::runtime_error here so a caller can recover.
#pragma once

#include <string>
#include <vector>

#include "gpv.h"
#include "gpv_mesh.h"

namespace gpv {

// The GLParameters fields the voxelization path reads
// (includes/GLParameters.h; defaults src/GLParameters.cpp:63-67).
struct GLParameters {
    int voxelCount = 8;
    int voxelCount2 = 4;
    bool level2Voxels = true;
    bool saveVoxels = true;
    std::string outputDir = ".";
    // the Level2 arithmetic of the reference's CUDACLASSIFYTESSELLATION build
    // (src/Object.cpp:3342-3347; gpv_params::cuda_level2)
    bool cudaClassifyTessellation = false;
    // Level1 inside/outside by the GL stencil fill the reference's INOUT build
    // runs (ClassifyInOut2x, src/Object.cpp:3156-3159; gpv_params::l1_stencil)
    // instead of the CPU composition's ray parity
    bool stencilInOut = false;
};

class Object {
public:
    explicit Object(gpv_ctx* ctx = nullptr);
    ~Object();
    Object(const Object&) = delete;
    Object& operator=(const Object&) = delete;

    void ReadObject(const char* fname);
    void ReadOFFObject(const char* fname);
    void CreateFlatTriangleData();
    void PerformVoxelization(GLParameters* glParam, int bufferSize);
    void SaveVoxelization(GLParameters* glParam);

    // host copies of the result (filled by PerformVoxelization)
    void Download(std::vector<uint8_t>* level1InOut, std::vector<int32_t>* boundaryPrefixSum,
                  std::vector<int32_t>* boundaryIndex, std::vector<uint8_t>* level2InOut) const;

    int objID = -1;
    float bBoxMin[3] = {0, 0, 0};
    float bBoxMax[3] = {0, 0, 0};
    float maxModelSize = 0.0f;
    int64_t totalNumTriangles = 0;
    std::vector<float> flatCPUTriangleData;
    gpv_grid_info voxelData{};          // dims, grid sizes, counts, timings
    gpv_grid* grid() const { return grid_; }
    gpv_ctx* context() const { return ctx_; }

private:
    void adopt(gpv_mesh& m);
    gpv_ctx* ctx_ = nullptr;
    bool own_ctx_ = false;
    gpv_mesh mesh_{};
    gpv_grid* grid_ = nullptr;
};

}  // namespace gpv
