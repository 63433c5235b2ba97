// Decodes a mesh file with the C++ decoders (include/bvhtracer_mesh.hpp) and
// prints every triangle's vertices, texture coordinates and normals as raw
// binary32 words (hex), one triangle per line: the form the tests compare
// with the Python decoders' output.  Usage: mesh_dump tri|obj FILE
#include <cstdio>
#include <cstring>
#include <fstream>

#include "bvhtracer_mesh.hpp"

using namespace bvhtracer;

static void word(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    std::printf(" %08x", u);
}

int main(int argc, char** argv)
{
    if (argc != 3)
        return 2;
    std::ifstream in(argv[2], std::ios::binary);
    if (!in)
        return 2;
    try {
        Mesh mesh = std::strcmp(argv[1], "tri") == 0 ? TriMeshDecoder(in).read_mesh() : ObjMeshDecoder(in).read_mesh();
        for (size_t i = 0; i < mesh.len(); ++i) {
            for (const Vector3& v : mesh.primitives()[i].vertices)
                word(v.x), word(v.y), word(v.z);
            for (const auto& uv : mesh.tex_coords()[i].uv)
                word(uv[0]), word(uv[1]);
            for (const Vector3& v : mesh.normals()[i].normals)
                word(v.x), word(v.y), word(v.z);
            std::printf("\n");
        }
    } catch (const TriLoaderError& e) {
        std::printf("TriLoaderError %zu %d %s\n", e.line_number(), (int)e.kind(), e.what());
        return 1;
    } catch (const Error& e) {
        std::printf("Error %s\n", e.what());
        return 1;
    }
    return 0;
}
