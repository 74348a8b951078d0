// Test-only driver of the Object-level drop-in (include/gpv_object.hpp): the
// sequence GPView runs per object on its voxelization key
// (src/GPView.cpp:1374-1381): ReadObject -> CreateFlatTriangleData ->
// PerformVoxelization(glParam, bufferSize) -> SaveVoxelization (called by
// PerformVoxelization when glParam->saveVoxels).
//
//   object_dropin <mesh.obj|mesh.off> <voxelCount> <voxelCount2> <bufferSize> <outdir> <objID>
//
// Prints one line of the grid's counts; the .raw files go to outdir.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "gpv_object.hpp"

int main(int argc, char** argv)
{
    if (argc != 7) {
        fprintf(stderr, "usage: %s mesh voxelCount voxelCount2 bufferSize outdir objID\n", argv[0]);
        return 2;
    }
    try {
        gpv::GLParameters glParam;
        glParam.voxelCount = atoi(argv[2]);
        glParam.voxelCount2 = atoi(argv[3]);
        glParam.level2Voxels = glParam.voxelCount2 > 0;
        glParam.saveVoxels = true;
        glParam.outputDir = argv[5];
        gpv::Object obj;
        obj.objID = atoi(argv[6]);
        const size_t n = strlen(argv[1]);
        if (n > 4 && strcmp(argv[1] + n - 4, ".off") == 0)
            obj.ReadOFFObject(argv[1]);
        else
            obj.ReadObject(argv[1]);
        obj.CreateFlatTriangleData();
        obj.PerformVoxelization(&glParam, atoi(argv[4]));
        const gpv_grid_info& v = obj.voxelData;
        printf("triangles %lld grid %d %d %d n2 %d boundary %lld inside %lld l2_inside %lld l2_boundary %lld "
               "regrows %d\n",
               (long long)obj.totalNumTriangles, v.n[0], v.n[1], v.n[2], v.n2, (long long)v.nb,
               (long long)v.l1_inside, (long long)v.l2_inside, (long long)v.l2_boundary, v.regrows);
    } catch (const std::exception& e) {
        fprintf(stderr, "object_dropin: %s\n", e.what());
        return 1;
    }
    return 0;
}
