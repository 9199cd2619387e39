// Host-side input formats: binary / ASCII STL
// (PolygonMesh::readPolygonMeshFile, src/PolygonMesh.cpp:42-108) and the .grid
// particle file (MpsInputOutput::readMpsParticleFile,
// src/MpsInputOutput.cpp:645-857).  VTU output is in output.hip.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "mps_gpu.h"
#include "mps_internal.h"

namespace {

struct Tri {
    float v[3][3];
};

bool readStlBinary(FILE* f, std::vector<Tri>& tris) {
    if (fseek(f, 80, SEEK_SET) != 0) return false;
    uint32_t n = 0;
    if (fread(&n, 4, 1, f) != 1) return false;
    fseek(f, 0, SEEK_END);
    long size = ftell(f);
    if (size != 84 + 50L * (long)n) return false;
    fseek(f, 84, SEEK_SET);
    tris.resize(n);
    for (uint32_t t = 0; t < n; t++) {
        float rec[12];
        uint16_t attr;
        if (fread(rec, 4, 12, f) != 12 || fread(&attr, 2, 1, f) != 1) return false;
        for (int v = 0; v < 3; v++)
            for (int c = 0; c < 3; c++) tris[t].v[v][c] = rec[3 + v * 3 + c];
    }
    return true;
}

bool readStlAscii(FILE* f, std::vector<Tri>& tris) {
    fseek(f, 0, SEEK_SET);
    char word[256];
    tris.clear();
    Tri cur;
    int nv = 0;
    while (fscanf(f, "%255s", word) == 1) {
        if (strcmp(word, "vertex") == 0) {
            float x, y, z;
            if (fscanf(f, "%f %f %f", &x, &y, &z) != 3) return false;
            if (nv < 3) { cur.v[nv][0] = x; cur.v[nv][1] = y; cur.v[nv][2] = z; }
            nv++;
        } else if (strcmp(word, "endfacet") == 0) {
            if (nv != 3) return false;
            tris.push_back(cur);
            nv = 0;
        }
    }
    return !tris.empty();
}

}  // namespace

extern "C" {

// PolygonMesh::writePolygonMeshFile (src/PolygonMesh.cpp:110-139) writes the
// mesh with igl::writeSTL; here an ASCII STL with each facet's normal from
// its vertices.
int mps_mesh_write_stl(const char* path, int nv, const double* V, int nf, const int* F) {
    if (!path || nv <= 0 || nf <= 0 || !V || !F) return MPS_E_ARG;
    FILE* fp = fopen(path, "w");
    if (!fp) {
        mpsgpu::setGlobalError(std::string("cannot write ") + path);
        return MPS_E_IO;
    }
    fprintf(fp, "solid mesh\n");
    for (int t = 0; t < nf; t++) {
        const double* a = &V[(size_t)F[t * 3] * 3];
        const double* b = &V[(size_t)F[t * 3 + 1] * 3];
        const double* c = &V[(size_t)F[t * 3 + 2] * 3];
        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double r = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; k++) n[k] = (r == 0) ? 0.0 : n[k] / r;
        fprintf(fp, "facet normal %e %e %e\nouter loop\n", n[0], n[1], n[2]);
        for (const double* v : {a, b, c}) fprintf(fp, "vertex %e %e %e\n", v[0], v[1], v[2]);
        fprintf(fp, "endloop\nendfacet\n");
    }
    fprintf(fp, "endsolid mesh\n");
    const bool ok = fclose(fp) == 0;
    (void)nv;
    return ok ? MPS_OK : MPS_E_IO;
}

int mps_mesh_read_stl(const char* path, int* nv_out, double* V, int* nf_out, int* F, double* N) {
    if (!path || !nv_out || !nf_out) return MPS_E_ARG;
    FILE* f = fopen(path, "rb");
    if (!f) { mpsgpu::setGlobalError(std::string("cannot open ") + path); return MPS_E_IO; }
    std::vector<Tri> tris;
    bool ok = readStlBinary(f, tris);
    if (!ok) ok = readStlAscii(f, tris);
    fclose(f);
    if (!ok) { mpsgpu::setGlobalError(std::string("cannot parse STL ") + path); return MPS_E_IO; }
    // merge duplicate vertices: STL coordinates are float32, so the 2e-15
    // tolerance of the reference merges exactly the identical ones
    std::map<std::tuple<float, float, float>, int> idx;
    std::vector<double> verts;
    std::vector<int> faces;
    for (const Tri& t : tris)
        for (int v = 0; v < 3; v++) {
            auto key = std::make_tuple(t.v[v][0], t.v[v][1], t.v[v][2]);
            auto it = idx.find(key);
            int id;
            if (it == idx.end()) {
                id = (int)(verts.size() / 3);
                idx[key] = id;
                for (int c = 0; c < 3; c++) verts.push_back((double)t.v[v][c]);
            } else {
                id = it->second;
            }
            faces.push_back(id);
        }
    const int nv = (int)(verts.size() / 3), nf = (int)tris.size();
    if (V && *nv_out >= nv) memcpy(V, verts.data(), sizeof(double) * verts.size());
    if (F && *nf_out >= nf) memcpy(F, faces.data(), sizeof(int) * faces.size());
    if (N && *nf_out >= nf) {
        // unit face normals (v1 - v0) x (v2 - v0); zero for degenerate faces
        for (int t = 0; t < nf; t++) {
            const double* a = &verts[faces[t * 3] * 3];
            const double* b = &verts[faces[t * 3 + 1] * 3];
            const double* c = &verts[faces[t * 3 + 2] * 3];
            const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            const double r = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int k = 0; k < 3; k++) N[t * 3 + k] = (r == 0) ? 0.0 : n[k] / r;
        }
    }
    *nv_out = nv;
    *nf_out = nf;
    return MPS_OK;
}

int mps_read_grid(const char* path, int* n_out, int* type, double* pos, double* vel, double* press,
                  double* press_avg) {
    if (!path || !n_out) return MPS_E_ARG;
    FILE* f = fopen(path, "r");
    if (!f) { mpsgpu::setGlobalError(std::string("cannot open ") + path); return MPS_E_IO; }
    int zero = 0, n = 0;
    if (fscanf(f, "%d", &zero) != 1 || fscanf(f, "%d", &n) != 1 || n < 0) {
        fclose(f);
        mpsgpu::setGlobalError("bad grid header");
        return MPS_E_IO;
    }
    if (!type) { fclose(f); *n_out = n; return MPS_OK; }
    if (*n_out < n) { fclose(f); mpsgpu::setGlobalError("grid arrays too small"); return MPS_E_ARG; }
    for (int i = 0; i < n; i++) {
        int t;
        double b[8];
        if (fscanf(f, "%d %lf %lf %lf %lf %lf %lf %lf %lf", &t, &b[0], &b[1], &b[2], &b[3], &b[4], &b[5], &b[6], &b[7]) != 9) {
            fclose(f);
            mpsgpu::setGlobalError("truncated grid file");
            return MPS_E_IO;
        }
        type[i] = t;
        for (int c = 0; c < 3; c++) { pos[i * 3 + c] = b[c]; vel[i * 3 + c] = b[3 + c]; }
        press[i] = b[6];
        press_avg[i] = b[7];
    }
    fclose(f);
    *n_out = n;
    return MPS_OK;
}

int mps_write_grid(const char* path, int n, const int* type, const double* pos, const double* vel,
                   const double* press, const double* press_avg) {
    if (!path || n < 0 || !type || !pos || !vel || !press || !press_avg) return MPS_E_ARG;
    FILE* f = fopen(path, "w");
    if (!f) { mpsgpu::setGlobalError(std::string("cannot write ") + path); return MPS_E_IO; }
    fprintf(f, "0\n%d\n", n);
    for (int i = 0; i < n; i++)
        fprintf(f, "%d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", type[i], pos[i * 3], pos[i * 3 + 1],
                pos[i * 3 + 2], vel[i * 3], vel[i * 3 + 1], vel[i * 3 + 2], press[i], press_avg[i]);
    fclose(f);
    return MPS_OK;
}

}  // extern "C"
