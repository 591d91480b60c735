// Near-field sampling planes: the -P argument and its mesh (src/sampling_planes.cpp:30-158), whose
// vertices eval_fields(mesh) samples (src/emw_postpro_common.cpp:77-90, write_fields :161-201).
//
// The reference meshes the plane with gmsh (an unstructured mesh of element size h on the
// rectangle), which is not available to this library; the plane here is a structured mesh of the
// same rectangle, ceil(width / h) x ceil(height / h) cells split into right triangles, placed and
// rotated as the reference places and rotates its rectangle.  The fields at its vertices come from
// frico_fields (one solution column, H with the reference's exp(+jkR)).
#include "../../include/frico_gpu.h"
#include "common.hpp"

#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

namespace frico {

// utils.cpp split(): tokens between separators, empty ones dropped (src/utils.cpp:30-55).
static std::vector<std::string> split_tokens(const char* s, char sep)
{
    std::vector<std::string> out;
    const std::string str(s);
    size_t first = 0;
    while (first < str.size()) {
        while (first < str.size() && str[first] == sep)
            first++;
        size_t last = first;
        while (last < str.size() && str[last] != sep)
            last++;
        if (last > first)
            out.push_back(str.substr(first, last - first));
        first = last + 1;
    }
    return out;
}

// std::stod: the longest numeric prefix after leading blanks; an error if there is none.
static bool to_double(const std::string& s, double* v)
{
    const char* b = s.c_str();
    char* e = nullptr;
    *v = std::strtod(b, &e);
    return e != b;
}

} // namespace frico

extern "C" {

int frico_sampling_plane_parse(const char* spec, double* axis, double* width, double* height, double* h)
{
    using namespace frico;
    if (!spec)
        return set_error("null plane specification"), -1;
    if (!axis || !width || !height || !h)
        return set_error("null output"), -2;
    const auto toks = split_tokens(spec, ':');
    double v[6];
    if (toks.size() == 4) {
        // short form plane:width:height:h (src/sampling_planes.cpp:47-82)
        const std::string& p = toks[0];
        if (p == "xy") {
            axis[0] = 0, axis[1] = 0, axis[2] = 1;
        } else if (p == "yz") {
            axis[0] = 0, axis[1] = 1, axis[2] = 0;
        } else if (p == "xz") {
            axis[0] = 1, axis[1] = 0, axis[2] = 0;
        } else {
            return set_error("sampling plane '%s': unknown plane '%s' (xy, yz or xz)", spec, p.c_str()), -1;
        }
        for (int i = 0; i < 3; i++)
            if (!to_double(toks[1 + i], &v[i]))
                return set_error("sampling plane '%s': '%s' is not a number", spec, toks[1 + i].c_str()), -1;
        *width = v[0], *height = v[1], *h = v[2];
        return 0;
    }
    if (toks.size() == 6) {
        // long form nx:ny:nz:width:height:h, the axis normalised (src/sampling_planes.cpp:84-106)
        for (int i = 0; i < 6; i++)
            if (!to_double(toks[i], &v[i]))
                return set_error("sampling plane '%s': '%s' is not a number", spec, toks[i].c_str()), -1;
        const double nn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (!(nn > 0))
            return set_error("sampling plane '%s': zero axis", spec), -1;
        axis[0] = v[0] / nn, axis[1] = v[1] / nn, axis[2] = v[2] / nn;
        *width = v[3], *height = v[4], *h = v[5];
        return 0;
    }
    return set_error("sampling plane '%s': expected plane:width:height:h or nx:ny:nz:width:height:h", spec), -1;
}

int frico_sampling_plane_mesh(const double* axis, double width, double height, double h, double* vertices,
                              int64_t max_vertices, int64_t* triangles, int64_t max_triangles, int64_t* nv_out,
                              int64_t* nt_out)
{
    using namespace frico;
    if (!axis)
        return set_error("null axis"), -1;
    if (!(width > 0) || !(height > 0))
        return set_error("plane width and height must be positive"), -2;
    if (!(h > 0))
        return set_error("plane element size must be positive"), -4;
    const double cx = std::ceil(width / h), cy = std::ceil(height / h);
    if (cx * cy > 1e8)
        return set_error("sampling plane of %.0f x %.0f cells is too large", cx, cy), -4;
    const int64_t nx = cx < 1 ? 1 : (int64_t)cx, ny = cy < 1 ? 1 : (int64_t)cy;
    const int64_t nv = (nx + 1) * (ny + 1), nt = 2 * nx * ny;
    if (nv_out)
        *nv_out = nv;
    if (nt_out)
        *nt_out = nt;
    if (!vertices && !triangles)
        return 0; // sizes only
    if (vertices && max_vertices < nv)
        return set_error("vertex buffer holds %lld, the plane has %lld", (long long)max_vertices, (long long)nv), -6;
    if (triangles && max_triangles < nt)
        return set_error("triangle buffer holds %lld, the plane has %lld", (long long)max_triangles, (long long)nt), -8;
    const double an = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    if (!(an > 0))
        return set_error("zero rotation axis"), -1;
    const double k[3] = {axis[0] / an, axis[1] / an, axis[2] / an};
    if (vertices) {
        // rectangle [-w/2, w/2] x [-h/2, h/2] in z = 0, turned by pi/2 about the axis through the
        // origin (src/sampling_planes.cpp:114-125): v' = k x v + k (k . v)  (Rodrigues at 90 deg)
        for (int64_t j = 0; j <= ny; j++)
            for (int64_t i = 0; i <= nx; i++) {
                const double x = -width / 2 + width * (double)i / (double)nx;
                const double y = -height / 2 + height * (double)j / (double)ny;
                const double kv = k[0] * x + k[1] * y;
                double* o = vertices + 3 * (j * (nx + 1) + i);
                o[0] = k[1] * 0.0 - k[2] * y + k[0] * kv;
                o[1] = k[2] * x - k[0] * 0.0 + k[1] * kv;
                o[2] = k[0] * y - k[1] * x + k[2] * kv;
            }
    }
    if (triangles) {
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                const int64_t v00 = j * (nx + 1) + i, v10 = v00 + 1, v01 = v00 + nx + 1, v11 = v01 + 1;
                int64_t* t = triangles + 6 * (j * nx + i);
                t[0] = v00, t[1] = v10, t[2] = v11;
                t[3] = v00, t[4] = v11, t[5] = v01;
            }
    }
    return 0;
}

} // extern "C"
