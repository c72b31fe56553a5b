// Host sanitizer driver of the OBJ loader (palz_mesh_load_obj and friends,
// palzilla_amd/csrc/scene_io.hip): loads every file named on the command line
// and copies its triangles out.  Built by tests/test_sanitizers.py with
// AddressSanitizer and UndefinedBehaviorSanitizer on the host code.
// Prints one line per file: "<path> <status> <triangles>".
#include <cstdio>
#include <vector>

#include "palz.h"

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        palz_mesh* m = nullptr;
        const int rc = palz_mesh_load_obj(argv[i], 0, &m);
        uint32_t n = 0;
        if (rc == PALZ_OK) {
            palz_mesh_triangle_count(m, &n);
            std::vector<palz_triangle> tris(n);
            std::vector<palz_triangle_attrib> attribs(n);
            palz_mesh_copy(m, tris.data(), attribs.data());
            palz_mesh_copy(m, nullptr, attribs.data());
            palz_mesh_destroy(m);
        }
        std::printf("%s %d %u\n", argv[i], rc, n);
    }
    return 0;
}
