// gpv_object.cpp -- gpv::Object facade (include/gpv_object.hpp).
#include "gpv_object.hpp"

#include <cstdio>
#include <cstdlib>
#include <stdexcept>

namespace gpv {

namespace {
[[noreturn]] void device_error(const char* what, int st)
{
    throw std::runtime_error(std::string(what) + ": " + gpv_status_string(st) + " (" + gpv_last_error() + ")");
}
}  // namespace

Object::Object(gpv_ctx* ctx) : ctx_(ctx) {}

Object::~Object()
{
    if (grid_) gpv_grid_release(ctx_, grid_);
    if (own_ctx_) gpv_ctx_destroy(ctx_);
    gpv_mesh_free(&mesh_);
}

void Object::adopt(gpv_mesh& m)
{
    gpv_mesh_free(&mesh_);
    mesh_ = m;
    if (gpv_mesh_bounds(&mesh_, bBoxMin, bBoxMax, &maxModelSize) != GPV_OK) {
        fprintf(stderr, "Object has no vertices\n");
        abort();
    }
}

void Object::ReadObject(const char* fname)
{
    gpv_mesh m{};
    const int st = gpv_mesh_read_obj(fname, &m);
    if (st == GPV_ERR_IO) {
        fprintf(stderr, "Unable to open file \"%s\"\n", fname);
        abort();
    }
    if (st != GPV_OK) {
        fprintf(stderr, "Unsupported OBJ content in \"%s\"\n", fname);
        abort();
    }
    adopt(m);
}

void Object::ReadOFFObject(const char* fname)
{
    gpv_mesh m{};
    const int st = gpv_mesh_read_off(fname, &m);
    if (st == GPV_ERR_IO) {
        fprintf(stderr, "Unable to open file \"%s\"\n", fname);
        abort();
    }
    if (st != GPV_OK) {
        fprintf(stderr, "Unsupported OFF content in \"%s\"\n", fname);
        abort();
    }
    adopt(m);
}

void Object::CreateFlatTriangleData()
{
    totalNumTriangles = mesh_.ntris;
    flatCPUTriangleData.assign((size_t)mesh_.ntris * 9, 0.0f);
    if (mesh_.ntris > 0) gpv_mesh_flatten(&mesh_, flatCPUTriangleData.data());
}

void Object::PerformVoxelization(GLParameters* glParam, int bufferSize)
{
    if (!ctx_) {
        int st = gpv_ctx_create(-1, 0, &ctx_);
        if (st != GPV_OK) device_error("gpv_ctx_create", st);
        own_ctx_ = true;
    }
    if (grid_) {
        gpv_grid_release(ctx_, grid_);
        grid_ = nullptr;
    }
    if ((int64_t)flatCPUTriangleData.size() != totalNumTriangles * 9) CreateFlatTriangleData();
    // the reference always computes the normals it saves (src/Object.cpp:3219-3253, 2404-2421)
    gpv_params p{glParam->voxelCount, glParam->voxelCount2, glParam->level2Voxels ? 1 : 0, 1,
                 bufferSize > 0 ? (int64_t)bufferSize : 0, glParam->cudaClassifyTessellation ? 1 : 0,
                 glParam->stencilInOut ? 1 : 0};
    int st = gpv_voxelize(ctx_, flatCPUTriangleData.data(), totalNumTriangles, 0, bBoxMin, bBoxMax, maxModelSize,
                          &p, &grid_);
    if (st != GPV_OK) device_error("PerformVoxelization", st);
    gpv_grid_get_info(grid_, &voxelData);
    if (glParam->saveVoxels) SaveVoxelization(glParam);
}

void Object::SaveVoxelization(GLParameters* glParam)
{
    if (!grid_) return;
    const int st = gpv_save_raw(ctx_, grid_, objID, glParam->outputDir.c_str());
    if (st == GPV_ERR_IO) {
        fprintf(stderr, "Unable to open output file for writing\n");
        abort();
    }
    if (st != GPV_OK) device_error("SaveVoxelization", st);
}

void Object::Download(std::vector<uint8_t>* l1, std::vector<int32_t>* prefix, std::vector<int32_t>* bidx,
                      std::vector<uint8_t>* l2) const
{
    if (!grid_) return;
    const int64_t n23 = (int64_t)voxelData.n2 * voxelData.n2 * voxelData.n2;
    if (l1) l1->assign((size_t)voxelData.n1, 0);
    if (prefix) prefix->assign((size_t)voxelData.n1, 0);
    if (bidx) bidx->assign((size_t)voxelData.nb, 0);
    if (l2) l2->assign((size_t)(voxelData.nb * n23), 0);
    const int st = gpv_grid_download(ctx_, grid_, l1 ? l1->data() : nullptr, prefix ? prefix->data() : nullptr,
                                     bidx ? bidx->data() : nullptr, l2 ? l2->data() : nullptr);
    if (st != GPV_OK) device_error("Download", st);
}

}  // namespace gpv
