// VTU particle output from device-resident state.
//
// ASCII: MpsInputOutput::writeVtuAscii (src/MpsInputOutput.cpp:1313-1579),
// byte for byte (same arrays, same printf conversions of the same values).
// Binary: the same arrays as a VTU with appended raw data and UInt64 block
// headers.  The reference's binary writer (src/MpsInputOutput.cpp:921-1108)
// is marked as not working in include/MpsInputOutput.h:85; this one writes a
// file that VTK readers accept.
//
// A snapshot is taken in stream order: gather kernels on the context stream
// put every array the file needs into a device staging buffer, in reference
// order and already in the type the file holds (positions FP64, since the
// ASCII file prints them with %lf; the other real fields as the FP32 values
// both files hold; integers as int32).  A second stream copies the staging
// buffer into pinned host memory, and a writer thread formats the text (split
// over a few threads, with the printf-exact fixed-point formatter of
// fmt_fixed.h) and writes the file while the context stream runs the next
// steps.  Two slots let one snapshot be copied while the previous file
// is still being written; a third submission waits for the oldest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "fmt_fixed.h"
#include "mps_gpu.h"
#include "mps_internal.h"

namespace mpsgpu {

namespace {

enum ArrKind { K_POS = 0, K_F32 = 1, K_I32 = 2 };

struct ArrSpec {
    const char* name;
    int kind;
    int comps;
    int field;     // first D (real) or I (integer) field index
    size_t off;    // byte offset in the staging buffer
};

struct Slot {
    void* dev = nullptr;
    void* host = nullptr;   // pinned
    size_t cap = 0;
    hipEvent_t copied = nullptr;
    std::thread writer;
    bool busy = false;
    int rc = MPS_OK;
    std::string err;
};

__global__ void kGatherPos(int n, const int* __restrict__ canon, const double* __restrict__ x,
                           const double* __restrict__ y, const double* __restrict__ z, double* __restrict__ dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = canon[k];
    if (c < 0 || c >= n) return;
    dst[3 * (size_t)c] = x[k];
    dst[3 * (size_t)c + 1] = y[k];
    dst[3 * (size_t)c + 2] = z[k];
}

__global__ void kGatherF32(int n, const int* __restrict__ canon, const double* __restrict__ s0,
                           const double* __restrict__ s1, const double* __restrict__ s2, float* __restrict__ dst,
                           int comps) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = canon[k];
    if (c < 0 || c >= n) return;
    if (comps == 1) {
        dst[c] = (float)s0[k];
    } else {
        dst[3 * (size_t)c] = (float)s0[k];
        dst[3 * (size_t)c + 1] = (float)s1[k];
        dst[3 * (size_t)c + 2] = (float)s2[k];
    }
}

__global__ void kGatherI32(int n, const int* __restrict__ canon, const int* __restrict__ s, int* __restrict__ dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = canon[k];
    if (c < 0 || c >= n) return;
    dst[c] = s[k];
}

// The arrays of writeVtuAscii (or, with MPS_VTU_ONLY_FREE_SURFACE,
// writeVtuAsciiFreeSurface, src/MpsInputOutput.cpp:1582-1875) for the given
// output flags, in file order.  Arrays without a name are staged for the
// free-surface selection only.
std::vector<ArrSpec> layout(int flags, int wallType, int N, size_t* bytes) {
    const bool fs = (flags & MPS_VTU_ONLY_FREE_SURFACE) != 0;
    std::vector<ArrSpec> a;
    a.push_back({"Position", K_POS, 3, DX, 0});
    a.push_back({"Velocity", K_F32, 3, DU, 0});
    a.push_back({"pressure", K_F32, 1, DPRESS, 0});
    a.push_back({"BC", K_I32, 1, IBC, 0});
    if (flags & MPS_VTU_PND) {
        a.push_back({"pnd", K_F32, 1, DPNDI, 0});
        a.push_back({"pndSmall", K_F32, 1, DPNDSMALL, 0});
    }
    if (flags & MPS_VTU_NEIGH) a.push_back({"nNeigh", K_I32, 1, INNEIGH, 0});
    if (flags & MPS_VTU_DEVIATION) {
        a.push_back({"npcdDeviation", K_F32, 3, DNPX, 0});
        a.push_back({"polygonNormal", K_F32, 3, DPNX, 0});
    }
    if (flags & MPS_VTU_CONCENTRATION) {
        a.push_back({"concentration", K_F32, 1, DCONC, 0});
        a.push_back({"GradCi", K_F32, 3, DGCX, 0});
    }
    if (flags & MPS_VTU_NON_NEWTONIAN) {   // src/MpsInputOutput.cpp:1463-1495, 1732-1775
        a.push_back({"RHO", K_F32, 1, DRHO, 0});
        a.push_back({"ConcVol", K_F32, 1, DCV, 0});
        a.push_back({"MEU", K_F32, 1, DMEU, 0});
        a.push_back({"MEUy", K_F32, 1, DMEUY, 0});
        a.push_back({"II", K_F32, 1, DII, 0});
        a.push_back({"PTYPE", K_I32, 1, IPTYPE, 0});
        a.push_back({"p_smooth", K_F32, 1, DPSMOOTH, 0});
        if (!fs) {   // the free-surface file ends the block at p_smooth
            a.push_back({"Inertia", K_F32, 1, DINERTIA, 0});
            a.push_back({"Normal_Stress", K_F32, 1, DPRHEO, 0});
            a.push_back({"VF", K_F32, 1, DVF, 0});
        }
    }
    if (flags & MPS_VTU_AUXILIAR) {
        // the free-surface file has no ParticleType array
        if (!fs) a.push_back({"ParticleType", K_I32, 1, ITYPE, 0});
        if (wallType == MPS_WALL_POLYGON) a.push_back({"nearMeshType", K_I32, 1, INEARMESH, 0});
    }
    if (fs) a.push_back({nullptr, K_I32, 1, ITYPE, 0});
    size_t off = 0;
    for (ArrSpec& s : a) {
        s.off = off;
        const size_t elem = s.kind == K_POS ? sizeof(double) : 4;
        off += (elem * s.comps * (size_t)N + 255) & ~(size_t)255;
    }
    *bytes = off;
    return a;
}

// Particles of the free-surface file: free-surface BC or wall type
// (src/MpsInputOutput.cpp:1598-1602).
std::vector<int> freeSurfaceSelection(const std::vector<ArrSpec>& arrs, const char* base, int n, int surfaceId,
                                      int wallId) {
    const int *bc = nullptr, *type = nullptr;
    for (const ArrSpec& s : arrs) {
        if (s.name && strcmp(s.name, "BC") == 0) bc = reinterpret_cast<const int*>(base + s.off);
        if (!s.name) type = reinterpret_cast<const int*>(base + s.off);
    }
    std::vector<int> sel;
    for (int i = 0; i < n; i++)
        if (bc[i] == surfaceId || type[i] == wallId) sel.push_back(i);
    return sel;
}

// Formats items [0, n) with fmt(i, buf) (which writes at most 1024
// characters and returns how many) in chunks on up to `threads` threads and
// writes them in order; `typical` is the expected characters per item.
template <typename F>
bool writeFormatted(FILE* fp, int n, int threads, size_t typical, F fmt) {
    const int chunk = 1 << 15;
    const int nChunks = (n + chunk - 1) / chunk;
    if (nChunks == 0) return true;
    std::vector<std::string> text(nChunks);
    std::atomic<int> next(0);
    auto work = [&]() {
        char buf[1024];
        for (int c = next++; c < nChunks; c = next++) {
            const int lo = c * chunk, hi = std::min(n, lo + chunk);
            std::string& s = text[c];
            s.reserve((size_t)(hi - lo) * typical);
            for (int i = lo; i < hi; i++) s.append(buf, (size_t)fmt(i, buf));
        }
    };
    const int t = std::max(1, std::min(threads, nChunks));
    std::vector<std::thread> pool;
    for (int k = 1; k < t; k++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    bool ok = true;
    for (const std::string& s : text) ok = ok && fwrite(s.data(), 1, s.size(), fp) == s.size();
    return ok;
}

// "%f %f %f " / "%f " / "%d " of one item
inline int fmt3(char* b, double x, double y, double z) {
    int k = fmtFixed6(b, x);
    b[k++] = ' ';
    k += fmtFixed6(b + k, y);
    b[k++] = ' ';
    k += fmtFixed6(b + k, z);
    b[k++] = ' ';
    return k;
}
inline int fmt1(char* b, double x) {
    const int k = fmtFixed6(b, x);
    b[k] = ' ';
    return k + 1;
}
inline int fmtI(char* b, int v) {
    const int k = fmtInt(b, v);
    b[k] = ' ';
    return k + 1;
}

// sel: the particles written (all n when null, m = n).
int writeAscii(const char* path, const std::vector<ArrSpec>& arrs, const char* base, int n, const int* sel, int m,
               int threads) {
    FILE* fp = fopen(path, "w");
    if (!fp) return MPS_E_IO;
    bool ok = true;
    fprintf(fp, "<?xml version='1.0' encoding='UTF-8'?>\n");
    fprintf(fp, "<VTKFile xmlns='VTK' byte_order='LittleEndian' version='0.1' type='UnstructuredGrid'>\n");
    fprintf(fp, "  <UnstructuredGrid>\n");
    fprintf(fp, "    <Piece NumberOfCells='%d' NumberOfPoints='%d'>\n", m, m);
    for (size_t k = 0; k < arrs.size() && ok; k++) {
        const ArrSpec& s = arrs[k];
        if (!s.name) continue;
        if (k == 0) fprintf(fp, "      <Points>\n");
        if (k == 1) fprintf(fp, "      <PointData>\n");
        const char* type = s.kind == K_I32 ? "Int32" : "Float32";
        if (s.comps == 3)
            fprintf(fp, "        <DataArray type='%s' Name='%s' NumberOfComponents='3' format='ascii'>\n", type, s.name);
        else
            fprintf(fp, "        <DataArray type='%s' Name='%s' format='ascii'>\n", type, s.name);
        // writeVtuAsciiFreeSurface prints polygonNormal for every particle
        // (src/MpsInputOutput.cpp:1690-1694); kept, so that its files match
        const bool every = sel && strcmp(s.name, "polygonNormal") == 0;
        const int* ix = every ? nullptr : sel;
        const int cnt = every ? n : m;
        auto at = [ix](int i) { return (size_t)(ix ? ix[i] : i); };
        const char* p = base + s.off;
        if (s.kind == K_POS) {
            fprintf(fp, "          ");
            const double* v = reinterpret_cast<const double*>(p);
            ok = writeFormatted(fp, cnt, threads, 30, [v, at](int i, char* b) {
                const size_t j = at(i);
                return fmt3(b, v[3 * j], v[3 * j + 1], v[3 * j + 2]);
            });
        } else if (s.kind == K_F32 && s.comps == 3) {
            const float* v = reinterpret_cast<const float*>(p);
            ok = writeFormatted(fp, cnt, threads, 30, [v, at](int i, char* b) {
                const size_t j = at(i);
                return fmt3(b, v[3 * j], v[3 * j + 1], v[3 * j + 2]);
            });
        } else if (s.kind == K_F32) {
            const float* v = reinterpret_cast<const float*>(p);
            ok = writeFormatted(fp, cnt, threads, 12, [v, at](int i, char* b) { return fmt1(b, v[at(i)]); });
        } else {
            const int* v = reinterpret_cast<const int*>(p);
            ok = writeFormatted(fp, cnt, threads, 4, [v, at](int i, char* b) { return fmtI(b, v[at(i)]); });
        }
        fprintf(fp, "\n        </DataArray>\n");
        if (k == 0) fprintf(fp, "      </Points>\n");
    }
    fprintf(fp, "      </PointData>\n");
    fprintf(fp, "      <Cells>\n");
    fprintf(fp, "        <DataArray type='Int32' Name='connectivity' format='ascii'>\n");
    ok = ok && writeFormatted(fp, m, threads, 8, [](int i, char* b) { return fmtI(b, i); });
    fprintf(fp, "\n        </DataArray>\n");
    fprintf(fp, "        <DataArray type='Int32' Name='offsets' format='ascii'>\n");
    ok = ok && writeFormatted(fp, m, threads, 8, [](int i, char* b) { return fmtI(b, i + 1); });
    fprintf(fp, "\n        </DataArray>\n");
    fprintf(fp, "        <DataArray type='UInt8' Name='types' format='ascii'>\n");
    ok = ok && writeFormatted(fp, m, threads, 2, [](int, char* b) { b[0] = '1'; b[1] = ' '; return 2; });
    fprintf(fp, "\n        </DataArray>\n");
    fprintf(fp, "      </Cells>\n");
    fprintf(fp, "    </Piece>\n");
    fprintf(fp, "  </UnstructuredGrid>\n");
    fprintf(fp, "</VTKFile>\n");
    ok = !ferror(fp) && ok;
    ok = fclose(fp) == 0 && ok;
    return ok ? MPS_OK : MPS_E_IO;
}

// Binary file of the same arrays; with a selection every array (also
// polygonNormal) holds the selected particles only.
int writeBinary(const char* path, const std::vector<ArrSpec>& arrs, const char* base, const int* sel, int m) {
    FILE* fp = fopen(path, "wb");
    if (!fp) return MPS_E_IO;
    const uint64_t um = (uint64_t)m;
    auto bytesOf = [&](const ArrSpec& s) { return um * s.comps * 4; };   // Position is written as Float32
    uint64_t off = 0;
    auto decl = [&](const char* attrs, uint64_t bytes) {
        fprintf(fp, "        <DataArray %s format='appended' offset='%llu'/>\n", attrs, (unsigned long long)off);
        off += sizeof(uint64_t) + bytes;
    };
    char attrs[160];
    fprintf(fp, "<?xml version='1.0'?>\n");
    fprintf(fp, "<VTKFile type='UnstructuredGrid' version='1.0' byte_order='LittleEndian' header_type='UInt64'>\n");
    fprintf(fp, "  <UnstructuredGrid>\n");
    fprintf(fp, "    <Piece NumberOfPoints='%d' NumberOfCells='%d'>\n", m, m);
    for (size_t k = 0; k < arrs.size(); k++) {
        const ArrSpec& s = arrs[k];
        if (!s.name) continue;
        if (k == 0) fprintf(fp, "      <Points>\n");
        if (k == 1) fprintf(fp, "      <PointData>\n");
        snprintf(attrs, sizeof attrs, "type='%s' Name='%s' NumberOfComponents='%d'", s.kind == K_I32 ? "Int32" : "Float32",
                 s.name, s.comps);
        decl(attrs, bytesOf(s));
        if (k == 0) fprintf(fp, "      </Points>\n");
    }
    fprintf(fp, "      </PointData>\n");
    fprintf(fp, "      <Cells>\n");
    decl("type='Int64' Name='connectivity'", um * 8);
    decl("type='Int64' Name='offsets'", um * 8);
    decl("type='UInt8' Name='types'", um);
    fprintf(fp, "      </Cells>\n");
    fprintf(fp, "    </Piece>\n");
    fprintf(fp, "  </UnstructuredGrid>\n");
    fprintf(fp, "  <AppendedData encoding='raw'>\n_");
    bool ok = true;
    auto block = [&](const void* data, uint64_t bytes) {
        ok = ok && fwrite(&bytes, sizeof bytes, 1, fp) == 1 && (bytes == 0 || fwrite(data, 1, bytes, fp) == bytes);
    };
    std::vector<char> tmp;
    for (const ArrSpec& s : arrs) {
        if (!s.name) continue;
        const char* p = base + s.off;
        if (s.kind == K_POS) {
            const double* v = reinterpret_cast<const double*>(p);
            tmp.resize(bytesOf(s));
            float* f = reinterpret_cast<float*>(tmp.data());
            for (int i = 0; i < m; i++) {
                const size_t j = sel ? (size_t)sel[i] : (size_t)i;
                for (int c = 0; c < 3; c++) f[3 * (size_t)i + c] = (float)v[3 * j + c];
            }
            block(tmp.data(), bytesOf(s));
        } else if (sel) {
            const size_t w = 4 * (size_t)s.comps;
            tmp.resize(bytesOf(s));
            for (int i = 0; i < m; i++) memcpy(tmp.data() + w * i, p + w * (size_t)sel[i], w);
            block(tmp.data(), bytesOf(s));
        } else {
            block(p, bytesOf(s));
        }
    }
    std::vector<int64_t> idx((size_t)m);
    for (int i = 0; i < m; i++) idx[i] = i;
    block(idx.data(), um * 8);
    for (int i = 0; i < m; i++) idx[i] = (int64_t)i + 1;
    block(idx.data(), um * 8);
    std::vector<uint8_t> types((size_t)m, 1);
    block(types.data(), um);
    fprintf(fp, "\n  </AppendedData>\n</VTKFile>\n");
    ok = !ferror(fp) && ok;
    ok = fclose(fp) == 0 && ok;
    return ok ? MPS_OK : MPS_E_IO;
}

int formatThreads() {
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hw ? hw : 1u, 8u));
}

}  // namespace

struct VtuQueue {
    hipStream_t copy = nullptr;
    hipEvent_t gathered = nullptr;
    Slot slot[2];
    int next = 0;
    int rc = MPS_OK;          // first error of a finished writer not yet reported
    std::string err;
};

static void joinSlot(VtuQueue* q, Slot& s) {
    if (!s.busy) return;
    s.writer.join();
    s.busy = false;
    if (s.rc != MPS_OK && q->rc == MPS_OK) {
        q->rc = s.rc;
        q->err = s.err;
    }
}

int vtuSubmit(Ctx* ctx, const char* path, int flags) {
    if (!path) return MPS_E_ARG;
    if (!ctx->vtu) ctx->vtu = new VtuQueue();
    VtuQueue* q = ctx->vtu;
    MPS_HIP(hipSetDevice(ctx->dev));
    if (!q->copy) MPS_HIP(hipStreamCreateWithFlags(&q->copy, hipStreamNonBlocking));
    if (!q->gathered) MPS_HIP(hipEventCreateWithFlags(&q->gathered, hipEventDisableTiming));
    Slot& s = q->slot[q->next];
    q->next ^= 1;
    joinSlot(q, s);
    const int N = ctx->N;
    const int n = ctx->nLiveHost;
    size_t bytes = 0;
    std::vector<ArrSpec> arrs = layout(flags, ctx->P.wall_type, N, &bytes);
    if (bytes > s.cap) {
        if (s.dev) MPS_HIP(hipFree(s.dev));
        if (s.host) MPS_HIP(hipHostFree(s.host));
        s.dev = s.host = nullptr;
        s.cap = 0;
        MPS_HIP(hipMalloc(&s.dev, bytes));
        MPS_HIP(hipHostMalloc(&s.host, bytes, hipHostMallocDefault));
        s.cap = bytes;
    }
    if (!s.copied) MPS_HIP(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
    const int g = gridFor(N, 256);
    char* dev = static_cast<char*>(s.dev);
    for (const ArrSpec& a : arrs) {
        if (N == 0) break;
        if (a.kind == K_POS)
            hipLaunchKernelGGL(kGatherPos, dim3(g), dim3(256), 0, ctx->stream, N, ctx->I[ICANON], ctx->D[a.field],
                               ctx->D[a.field + 1], ctx->D[a.field + 2], reinterpret_cast<double*>(dev + a.off));
        else if (a.kind == K_F32)
            hipLaunchKernelGGL(kGatherF32, dim3(g), dim3(256), 0, ctx->stream, N, ctx->I[ICANON], ctx->D[a.field],
                               ctx->D[a.comps == 3 ? a.field + 1 : a.field], ctx->D[a.comps == 3 ? a.field + 2 : a.field],
                               reinterpret_cast<float*>(dev + a.off), a.comps);
        else
            hipLaunchKernelGGL(kGatherI32, dim3(g), dim3(256), 0, ctx->stream, N, ctx->I[ICANON], ctx->I[a.field],
                               reinterpret_cast<int*>(dev + a.off));
    }
    MPS_HIP(hipGetLastError());
    MPS_HIP(hipEventRecord(q->gathered, ctx->stream));
    MPS_HIP(hipStreamWaitEvent(q->copy, q->gathered, 0));
    MPS_HIP(hipMemcpyAsync(s.host, s.dev, bytes, hipMemcpyDeviceToHost, q->copy));
    MPS_HIP(hipEventRecord(s.copied, q->copy));
    s.busy = true;
    s.rc = MPS_OK;
    s.err.clear();
    const std::string file(path);
    const bool binary = (flags & MPS_VTU_BINARY) != 0;
    const bool freeSurface = (flags & MPS_VTU_ONLY_FREE_SURFACE) != 0;
    const int surfaceId = ctx->P.surface, wallId = ctx->P.wall;
    const int dev_id = ctx->dev;
    Slot* sp = &s;
    s.writer = std::thread([sp, arrs, file, binary, freeSurface, surfaceId, wallId, n, dev_id]() {
        hipSetDevice(dev_id);
        hipError_t e = hipEventSynchronize(sp->copied);
        if (e != hipSuccess) {
            sp->rc = MPS_E_HIP;
            sp->err = std::string("VTU snapshot copy: ") + hipGetErrorString(e);
            return;
        }
        const char* base = static_cast<const char*>(sp->host);
        std::vector<int> sel;
        if (freeSurface) sel = freeSurfaceSelection(arrs, base, n, surfaceId, wallId);
        const int* sp_sel = freeSurface ? sel.data() : nullptr;
        const int m = freeSurface ? (int)sel.size() : n;
        sp->rc = binary ? writeBinary(file.c_str(), arrs, base, sp_sel, m)
                        : writeAscii(file.c_str(), arrs, base, n, sp_sel, m, formatThreads());
        if (sp->rc != MPS_OK) sp->err = "cannot write " + file;
    });
    return MPS_OK;
}

int vtuWait(Ctx* ctx) {
    VtuQueue* q = ctx->vtu;
    if (!q) return MPS_OK;
    // oldest first
    joinSlot(q, q->slot[q->next]);
    joinSlot(q, q->slot[q->next ^ 1]);
    const int rc = q->rc;
    if (rc != MPS_OK) ctx->err = q->err;
    q->rc = MPS_OK;
    q->err.clear();
    return rc;
}

void vtuFree(Ctx* ctx) {
    VtuQueue* q = ctx->vtu;
    if (!q) return;
    vtuWait(ctx);
    for (Slot& s : q->slot) {
        if (s.dev) hipFree(s.dev);
        if (s.host) hipHostFree(s.host);
        if (s.copied) hipEventDestroy(s.copied);
    }
    if (q->gathered) hipEventDestroy(q->gathered);
    if (q->copy) hipStreamDestroy(q->copy);
    delete q;
    ctx->vtu = nullptr;
}

}  // namespace mpsgpu
