// Cell and face metrics of a mesh (src/mesh/geometry.cpp:24-150, 301-539),
// shared by the host (compute_geometry, mesh.cpp) and the device
// (k_cell_geometry / k_face_geometry, kernels_mesh.hip): one source for the
// arithmetic, so both give the same values bit for bit.
#pragma once

#include <cmath>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define CFDG_HD __host__ __device__
#else
#define CFDG_HD
#endif

namespace cfdg {
namespace geom {

struct CellMetrics {
    double vol, cx, cy, cz;
};

// Volume and centroid of a tetrahedron (type 0, 4 nodes) or hexahedron
// (type 3, 8 nodes) by decomposition into tetrahedra spanned by the vertex
// average x0 and each (sub-)triangle of the faces (geometry.cpp:24-150).  The
// face tables are those of the CGNS elements (cgnstables.hpp:45-75): local
// node lists whose right-hand-rule normal points out of the element.
CFDG_HD inline CellMetrics cell_metrics(int type, const double* x, const double* y, const double* z) {
    const int tet_faces[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
    const int hex_faces[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};
    const int nn = (type == 0) ? 4 : 8;
    const int nf = (type == 0) ? 4 : 6;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < nn; ++i) {
        sx += x[i];
        sy += y[i];
        sz += z[i];
    }
    const double inv_n = 1.0 / static_cast<double>(nn);
    const double x0 = sx * inv_n, y0 = sy * inv_n, z0 = sz * inv_n;

    double vol = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    for (int f = 0; f < nf; ++f) {
        if (type == 0) {
            const int a = tet_faces[f][0], b = tet_faces[f][1], c = tet_faces[f][2];
            const double e1x = x[b] - x[a], e1y = y[b] - y[a], e1z = z[b] - z[a];
            const double e2x = x[c] - x[a], e2y = y[c] - y[a], e2z = z[c] - z[a];
            const double Sx = 0.5 * (e1y * e2z - e1z * e2y);
            const double Sy = 0.5 * (e1z * e2x - e1x * e2z);
            const double Sz = 0.5 * (e1x * e2y - e1y * e2x);
            const double vt = (1.0 / 3.0) * (Sx * (x[a] - x0) + Sy * (y[a] - y0) + Sz * (z[a] - z0));
            vol += vt;
            mx += vt * (0.25 * (x0 + x[a] + x[b] + x[c]));
            my += vt * (0.25 * (y0 + y[a] + y[b] + y[c]));
            mz += vt * (0.25 * (z0 + z[a] + z[b] + z[c]));
        } else {
            double fx = 0.0, fy = 0.0, fz = 0.0;
            for (int j = 0; j < 4; ++j) {
                fx += x[hex_faces[f][j]];
                fy += y[hex_faces[f][j]];
                fz += z[hex_faces[f][j]];
            }
            fx *= 0.25;
            fy *= 0.25;
            fz *= 0.25;
            for (int j = 0; j < 4; ++j) {
                const int a = hex_faces[f][j], b = hex_faces[f][(j + 1) & 3];
                const double e1x = x[a] - fx, e1y = y[a] - fy, e1z = z[a] - fz;
                const double e2x = x[b] - fx, e2y = y[b] - fy, e2z = z[b] - fz;
                const double Sx = 0.5 * (e1y * e2z - e1z * e2y);
                const double Sy = 0.5 * (e1z * e2x - e1x * e2z);
                const double Sz = 0.5 * (e1x * e2y - e1y * e2x);
                const double vt = (1.0 / 3.0) * (Sx * (fx - x0) + Sy * (fy - y0) + Sz * (fz - z0));
                vol += vt;
                mx += vt * (0.25 * (x0 + fx + x[a] + x[b]));
                my += vt * (0.25 * (y0 + fy + y[a] + y[b]));
                mz += vt * (0.25 * (z0 + fz + z[a] + z[b]));
            }
        }
    }
    if (vol > 1e-15) {
        const double iv = 1.0 / vol;
        return {vol, mx * iv, my * iv, mz * iv};
    }
    return {vol, x0, y0, z0};
}

struct FaceMetrics {
    double cx, cy, cz, nx, ny, nz, area;
};

// Area vector, unit normal and centroid of a face with nn nodes (node k at
// (x[k], y[k], z[k])): a triangle directly, a polygon by the fan of
// sub-triangles from its vertex average, with the area-weighted centroid of
// the sub-triangles (geometry.cpp:376-470).  The normal is left unnormalised
// (area 0) when the area is not above 1e-15; callers reject such faces.
CFDG_HD inline FaceMetrics face_metrics(int nn, const double* x, const double* y, const double* z) {
    double Sx = 0.0, Sy = 0.0, Sz = 0.0, fcx = 0.0, fcy = 0.0, fcz = 0.0;
    if (nn == 3) {
        const double e1x = x[1] - x[0], e1y = y[1] - y[0], e1z = z[1] - z[0];
        const double e2x = x[2] - x[0], e2y = y[2] - y[0], e2z = z[2] - z[0];
        Sx = 0.5 * (e1y * e2z - e1z * e2y);
        Sy = 0.5 * (e1z * e2x - e1x * e2z);
        Sz = 0.5 * (e1x * e2y - e1y * e2x);
        fcx = (x[0] + x[1] + x[2]) / 3.0;
        fcy = (y[0] + y[1] + y[2]) / 3.0;
        fcz = (z[0] + z[1] + z[2]) / 3.0;
    } else {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int k = 0; k < nn; ++k) {
            sx += x[k];
            sy += y[k];
            sz += z[k];
        }
        const double inv = 1.0 / static_cast<double>(nn);
        const double f0x = sx * inv, f0y = sy * inv, f0z = sz * inv;
        double wx = 0.0, wy = 0.0, wz = 0.0, wa = 0.0;
        for (int k = 0; k < nn; ++k) {
            const int p = k, q = (k + 1 == nn) ? 0 : k + 1;
            const double e1x = x[p] - f0x, e1y = y[p] - f0y, e1z = z[p] - f0z;
            const double e2x = x[q] - f0x, e2y = y[q] - f0y, e2z = z[q] - f0z;
            const double tx = 0.5 * (e1y * e2z - e1z * e2y);
            const double ty = 0.5 * (e1z * e2x - e1x * e2z);
            const double tz = 0.5 * (e1x * e2y - e1y * e2x);
            const double ta = sqrt(tx * tx + ty * ty + tz * tz);
            Sx += tx;
            Sy += ty;
            Sz += tz;
            wx += ta * ((f0x + x[p] + x[q]) / 3.0);
            wy += ta * ((f0y + y[p] + y[q]) / 3.0);
            wz += ta * ((f0z + z[p] + z[q]) / 3.0);
            wa += ta;
        }
        if (wa > 1e-15) {
            const double ia = 1.0 / wa;
            fcx = wx * ia;
            fcy = wy * ia;
            fcz = wz * ia;
        } else {
            fcx = f0x;
            fcy = f0y;
            fcz = f0z;
        }
    }
    const double area = sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
    FaceMetrics r{fcx, fcy, fcz, Sx, Sy, Sz, area};
    if (area > 1e-15) {
        const double ia = 1.0 / area;
        r.nx = Sx * ia;
        r.ny = Sy * ia;
        r.nz = Sz * ia;
    }
    return r;
}

}  // namespace geom
}  // namespace cfdg
