// libtinymd_amd: the reference's md_* runtime ABI (md_abi.h) and this
// library's tmd_* extensions (tmd_ext.h) over one device-resident
// simulation on one GPU.
//
// State lives on the device as double4 arrays (x, y, z, pad) for positions,
// velocities and forces; ghosts occupy the slots after the locals, as in the
// reference's grid (core/grid.impala:267-277).  The host keeps AoS mirrors
// that md_copy_data_{to,from}_accelerator move, as the reference's
// ArrayData host mirrors do (core/layouts.impala:1-49).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "tmd_internal.hpp"
#include "../../include/md_abi.h"
#include "../../include/tmd_ext.h"

namespace tmd {

// ---------------------------------------------------------------------------
// errors
static std::string g_error;
static bool g_has_error = false;

void set_error(const char* file, int line, const char* what) {
    if (g_has_error) return;
    g_has_error = true;
    char buf[512];
    const char* base = strrchr(file, '/');
    snprintf(buf, sizeof buf, "tinymd_amd: %s (%s:%d)", what, base ? base + 1 : file, line);
    g_error = buf;
    fprintf(stderr, "%s\n", buf);
    fflush(stderr);
}

bool has_error() { return g_has_error; }

static int g_alloc_free_depth = 0;
AllocFree::AllocFree() { ++g_alloc_free_depth; }
AllocFree::~AllocFree() { --g_alloc_free_depth; }
bool alloc_forbidden() { return g_alloc_free_depth > 0; }

// ---------------------------------------------------------------------------
// device buffers
void DevBuf::reserve(size_t need, bool preserve, hipStream_t s) {
    if (need <= bytes || has_error()) return;
    if (alloc_forbidden()) {
        set_error(__FILE__, __LINE__, "allocation-free call: a device buffer would have to grow (set the state up "
                                      "with the allocating entry points first)");
        return;
    }
    void* np = nullptr;
    TMD_HIP(hipMalloc(&np, need));
    if (has_error()) return;
    if (preserve && p && bytes) TMD_HIP(hipMemcpyAsync(np, p, bytes, hipMemcpyDeviceToDevice, s));
    // work already queued on the stream may still read the old buffer
    if (p) {
        TMD_HIP(hipStreamSynchronize(s));
        TMD_HIP(hipFree(p));
    }
    p = np;
    bytes = need;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

// ---------------------------------------------------------------------------
// kernel timer (HIP events on the library stream)
void KernelTimer::begin(int, hipStream_t s, hipEvent_t* ev0) {
    *ev0 = nullptr;
    if (!enabled) return;
    TMD_HIP(hipEventCreate(ev0));
    TMD_HIP(hipEventRecord(*ev0, s));
}

void KernelTimer::end(int id, hipStream_t s, hipEvent_t ev0) {
    if (!enabled || !ev0) return;
    hipEvent_t ev1 = nullptr;
    TMD_HIP(hipEventCreate(&ev1));
    TMD_HIP(hipEventRecord(ev1, s));
    pending.push_back({id, {ev0, ev1}});
    if (pending.size() > 4096) collect();
}

void KernelTimer::collect() {
    for (auto& pe : pending) {
        TMD_HIP(hipEventSynchronize(pe.second.second));
        float ms = 0.f;
        TMD_HIP(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
        total_ms[pe.first] += ms;
        count[pe.first] += 1;
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    pending.clear();
}

void KernelTimer::reset() {
    collect();
    for (int k = 0; k < K_NKERNELS; ++k) {
        total_ms[k] = 0.0;
        count[k] = 0;
    }
}

// ---------------------------------------------------------------------------
static Sim* g_sim = nullptr;
Sim& sim() {
    if (!g_sim) g_sim = new Sim();
    return *g_sim;
}
bool sim_alive() { return g_sim != nullptr && g_sim->stream != nullptr; }

static bool ensure_stream(Sim& s) {
    if (s.stream) return true;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        set_error(__FILE__, __LINE__, "no HIP device available");
        return false;
    }
    TMD_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    if (has_error()) return false;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) == hipSuccess && lds > 0) s.lds_limit = lds;
    s.counters.reserve(64 * sizeof(int), false, s.stream);
    TMD_HIP(hipMemsetAsync(s.counters.p, 0, 64 * sizeof(int), s.stream));
    return !has_error();
}

static void finish(Sim& s) {
    if (s.sync_calls && s.stream && !has_error()) TMD_HIP(hipStreamSynchronize(s.stream));
}

static void alloc_atoms(Sim& s, int cap) {
    s.pos.reserve(sizeof(double4) * (size_t)cap, true, s.stream);
    s.vel.reserve(sizeof(double4) * (size_t)cap, true, s.stream);
    s.mass.reserve(sizeof(double) * (size_t)cap, true, s.stream);
    s.g_owner.reserve(sizeof(int) * (size_t)cap, true, s.stream);
    s.g_src.reserve(sizeof(int) * (size_t)cap, true, s.stream);
    s.g_code.reserve(sizeof(int) * (size_t)cap, true, s.stream);
    s.cell_of.reserve(sizeof(int) * (size_t)cap, false, s.stream);
    s.cell_slot.reserve(sizeof(int) * (size_t)cap, false, s.stream);
    s.cell_parts.reserve(sizeof(int) * (size_t)cap, false, s.stream);
    s.cell_inv.reserve(sizeof(int) * (size_t)cap, false, s.stream);
    s.inv_valid = 0;
    s.force.reserve(sizeof(double4) * (size_t)std::max(s.n, 1), false, s.stream);
    if (cap > s.cap_all) s.cap_all = cap;
}

static void upload(Sim& s) {
    // the host mirror holds the locals and, after a download, the ghosts
    const int nall = std::min<int>(s.n + s.ng, (int)std::min(s.h_pos.size(), s.h_vel.size()) / 3);
    if (nall < s.n || (int)s.h_mass.size() < nall || (int)s.h_force.size() < 3 * s.n) {
        set_error(__FILE__, __LINE__, "host mirror smaller than the number of local atoms");
        return;
    }
    if (nall == 0) return;
    alloc_atoms(s, std::max(s.cap_all, nall));
    std::vector<double4> p(nall), v(nall);
    for (int i = 0; i < nall; ++i) {
        p[i] = make_double4(s.h_pos[3 * i], s.h_pos[3 * i + 1], s.h_pos[3 * i + 2], 0.0);
        v[i] = make_double4(s.h_vel[3 * i], s.h_vel[3 * i + 1], s.h_vel[3 * i + 2], 0.0);
    }
    mirror_invalidate(s);
    s.pos_stale = 0;  // pos and vel are uploaded whole
    s.vel_mirrored = 0;
    s.vel_mirror_wanted = 0;
    TMD_HIP(hipMemcpyAsync(s.pos.p, p.data(), sizeof(double4) * nall, hipMemcpyHostToDevice, s.stream));
    TMD_HIP(hipMemcpyAsync(s.vel.p, v.data(), sizeof(double4) * nall, hipMemcpyHostToDevice, s.stream));
    TMD_HIP(hipMemcpyAsync(s.mass.p, s.h_mass.data(), sizeof(double) * nall, hipMemcpyHostToDevice, s.stream));
    std::vector<double4> f(s.n);
    for (int i = 0; i < s.n; ++i) f[i] = make_double4(s.h_force[3 * i], s.h_force[3 * i + 1], s.h_force[3 * i + 2], 0.0);
    TMD_HIP(hipMemcpyAsync(s.force.p, f.data(), sizeof(double4) * s.n, hipMemcpyHostToDevice, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    s.dev_valid = true;
}

static void download(Sim& s) {
    const int nall = s.n + s.ng;
    s.h_pos.resize(3 * (size_t)nall);
    s.h_vel.resize(3 * (size_t)nall);
    s.h_mass.resize(nall);
    s.h_force.assign(3 * (size_t)nall, 0.0);
    if (nall == 0 || !s.dev_valid) return;
    std::vector<double4> p(nall), v(nall), f(std::max(s.n, 1));
    TMD_HIP(hipMemcpyAsync(p.data(), s.pos.p, sizeof(double4) * nall, hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipMemcpyAsync(v.data(), s.vel.p, sizeof(double4) * nall, hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipMemcpyAsync(s.h_mass.data(), s.mass.p, sizeof(double) * nall, hipMemcpyDeviceToHost, s.stream));
    if (s.n) TMD_HIP(hipMemcpyAsync(f.data(), s.force.p, sizeof(double4) * s.n, hipMemcpyDeviceToHost, s.stream));
    TMD_HIP(hipStreamSynchronize(s.stream));
    if (has_error()) return;
    for (int i = 0; i < nall; ++i) {
        s.h_pos[3 * i] = p[i].x; s.h_pos[3 * i + 1] = p[i].y; s.h_pos[3 * i + 2] = p[i].z;
        s.h_vel[3 * i] = v[i].x; s.h_vel[3 * i + 1] = v[i].y; s.h_vel[3 * i + 2] = v[i].z;
    }
    for (int i = 0; i < s.n; ++i) {
        s.h_force[3 * i] = f[i].x; s.h_force[3 * i + 1] = f[i].y; s.h_force[3 * i + 2] = f[i].z;
    }
}

// device operations need the atoms on the device; the reference requires an
// explicit md_copy_data_to_accelerator first (main.cpp:677), this library
// uploads on first use if the caller skipped it.
static bool ready(Sim& s) {
    if (has_error()) return false;
    if (!ensure_stream(s)) return false;
    if (!s.dev_valid) upload(s);
    return !has_error();
}

static void set_geometry(Sim& s, const double* aabb, const double* rank_aabb, double cell_spacing) {
    make_geom(s.g, aabb, rank_aabb, cell_spacing);
}

void make_geom(Geom& G, const double* aabb, const double* rank_aabb, double cell_spacing) {
    for (int d = 0; d < 6; ++d) {
        G.world[d] = aabb[d];
        G.rank[d] = rank_aabb[d];
    }
    for (int d = 0; d < 3; ++d) {
        G.aabb[2 * d] = rank_aabb[2 * d] - cell_spacing;
        G.aabb[2 * d + 1] = rank_aabb[2 * d + 1] + cell_spacing;
        G.len[d] = aabb[2 * d + 1] - aabb[2 * d];
    }
    G.spacing = cell_spacing;
    // core/grid.impala:95-98
    G.nx = (int)std::floor((G.aabb[1] - G.aabb[0] + (cell_spacing / 2.0)) / cell_spacing) + 1;
    G.ny = (int)std::floor((G.aabb[3] - G.aabb[2] + (cell_spacing / 2.0)) / cell_spacing) + 1;
    G.nz = (int)std::floor((G.aabb[5] - G.aabb[4] + (cell_spacing / 2.0)) / cell_spacing) + 1;
    G.ncells = G.nx * G.ny * G.nz;
}

static double host_temperature(const Sim& s) {
    // core/thermo.impala:1-16 on the host mirror
    const double t_scale = 1.0 / (double)(s.n * 3 - 3);
    double t = 0.0;
    for (int i = 0; i < s.n; ++i) {
        const double* v = &s.h_vel[3 * i];
        t += (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * 1.0;
    }
    return t * t_scale;
}

static void rebuild(Sim& s, int half, double cutoff) {
    s.ng = 0;  // exchange_particles on one rank (comm/comm.impala:392-396)
    hipEvent_t ev = nullptr;
    if (s.sort_particles) {
        s.timer.begin(K_BIN, s.stream, &ev);
        sort_locals(s);  // wraps, then renumbers the locals in cell order
        s.timer.end(K_BIN, s.stream, ev);
    }
    s.timer.begin(K_BORDERS, s.stream, &ev);
    build_ghosts(s, true);  // md_enforce_pbc + borders
    s.timer.end(K_BORDERS, s.stream, ev);
    s.timer.begin(K_BIN, s.stream, &ev);
    s.defer_window_sync = 1;  // the list build reads the window statistics
    bin_particles(s);
    s.timer.end(K_BIN, s.stream, ev);
    s.timer.begin(K_NEIGHBOR, s.stream, &ev);
    build_neighbor_lists(s, half, cutoff);
    s.timer.end(K_NEIGHBOR, s.stream, ev);
    s.defer_window_sync = 0;
    finish_windows(s);  // if the list build stopped early
}

// ---------------------------------------------------------------------------
// Park-Miller minimal standard (Schrage), as initialize.h:29-49
static double pm_random(int* idum) {
    const int ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
    const double am = 1.0 / im;
    int k = (*idum) / iq;
    *idum = ia * (*idum - k * iq) - ir * k;
    if (*idum < 0) *idum += im;
    return am * (*idum);
}

// counter-based generator for the granular gas (SplitMix64 finaliser)
static unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------
// Scratch memory of the allocation-free thermo entry point
// (tmd_thermo_noalloc): the per-block partial sums of the kinetic and
// energy/virial reductions and their totals, in a fixed-size device array
// that every GPU the library runs on holds until the process ends.  Host
// code never uses the array's name as an address: its device address is
// looked up once per device (hipGetSymbolAddress, with that device current)
// and kept in g_scratch.  Every reduction writes all the partials it reads,
// so nothing needs clearing between calls; calls that overlap on one GPU
// race on the array.
constexpr size_t kScratchDoubles = size_t(1) << 17;  // 1 MiB
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];
static double* g_scratch[kMaxScratchDevices] = {};

static double* scratch_ptr() {
    int dev = -1;
    TMD_HIP(hipGetDevice(&dev));
    if (has_error()) return nullptr;
    if (dev < 0 || dev >= kMaxScratchDevices) {
        set_error(__FILE__, __LINE__, "allocation-free thermo: device index beyond the scratch table");
        return nullptr;
    }
    if (!g_scratch[dev]) {
        void* p = nullptr;
        TMD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (has_error()) return nullptr;
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

}  // namespace tmd

using namespace tmd;

// ===========================================================================
// reference ABI
extern "C" {

int md_initialize_grid(double const* masses, struct Vector3D const* positions, struct Vector3D const* velocities, int nparticles,
                       double const* aabb, double const* rank_aabb, double cell_spacing, int cell_capacity,
                       int neighborlist_capacity) {
    (void)cell_capacity;  // cells are CSR (counting sort): no per-cell capacity
    Sim& s = sim();
    if (has_error()) return 0;
    if (!ensure_stream(s)) return 0;
    if (nparticles < 0 || cell_spacing <= 0.0 || !(cell_spacing == cell_spacing)) {
        set_error(__FILE__, __LINE__, "md_initialize_grid: invalid particle count or cell spacing");
        return 0;
    }
    set_geometry(s, aabb, rank_aabb, cell_spacing);
    for (int d = 0; d < 3; ++d) {
        if (!(s.g.len[d] > 2.0 * cell_spacing)) {
            set_error(__FILE__, __LINE__, "md_initialize_grid: box length must exceed two cell spacings in every dimension");
            return 0;
        }
    }
    if ((long long)s.g.nx * s.g.ny * s.g.nz > 0x7fffffffLL) {
        set_error(__FILE__, __LINE__, "md_initialize_grid: cell grid too large");
        return 0;
    }
    s.n = nparticles;
    s.ng = 0;
    s.nb_cap = neighborlist_capacity > 0 ? neighborlist_capacity : 1;
    s.nb_valid = 0;
    s.windows_valid = 0;
    s.inv_valid = 0;
    mirror_invalidate(s);
    s.pos_stale = 0;
    s.vel_mirrored = 0;
    s.vel_mirror_wanted = 0;
    s.ghost_pos_stale = 0;
    s.dem_mirrored = 0;
    s.gsrc_valid = 0;
    s.h_mass.assign(masses, masses + nparticles);
    s.h_pos.resize(3 * (size_t)nparticles);
    s.h_vel.resize(3 * (size_t)nparticles);
    s.h_force.assign(3 * (size_t)nparticles, 0.0);
    for (int i = 0; i < nparticles; ++i) {
        s.h_pos[3 * i] = positions[i].x; s.h_pos[3 * i + 1] = positions[i].y; s.h_pos[3 * i + 2] = positions[i].z;
        s.h_vel[3 * i] = velocities[i].x; s.h_vel[3 * i + 1] = velocities[i].y; s.h_vel[3 * i + 2] = velocities[i].z;
    }
    s.dev_valid = false;
    // ghosts of a cube of side L with a shell of width s: ~ N ((1 + 2s/L)^3 - 1)
    double shell = 1.0;
    for (int d = 0; d < 3; ++d) shell *= 1.0 + 2.0 * cell_spacing / s.g.len[d];
    const long long est = (long long)nparticles + (long long)(nparticles * (shell - 1.0) * 1.25) + 4096;
    alloc_atoms(s, (int)std::min<long long>(est, 0x7fffffffLL));
    return has_error() ? 0 : nparticles;
}

void md_rescale_grid(double const* aabb_data) {
    // core/runtime.impala:39-50: new rank AABB extended by one spacing
    Sim& s = sim();
    if (has_error()) return;
    double ra[6];
    for (int d = 0; d < 6; ++d) ra[d] = aabb_data[d];
    set_geometry(s, s.g.world, ra, s.g.spacing);
}

void md_copy_data_to_accelerator() {
    Sim& s = sim();
    if (has_error() || !ensure_stream(s)) return;
    upload(s);
}

void md_copy_data_from_accelerator() {
    Sim& s = sim();
    if (has_error() || !s.stream) return;
    download(s);
}

void md_initial_integration(double dt) {
    Sim& s = sim();
    if (!ready(s)) return;
    initial_integration(s, dt);
    s.integration_iters = s.n;
    finish(s);
}

void md_final_integration(double dt) {
    Sim& s = sim();
    if (!ready(s)) return;
    final_integration(s, dt);
    s.integration_iters = s.n;
    finish(s);
}

void md_enforce_pbc() {
    Sim& s = sim();
    if (!ready(s)) return;
    if (s.sort_particles) {
        sort_locals(s);  // the wrap, then the locals in cell order (the ghosts are rebuilt next)
        s.ng = 0;
    } else {
        enforce_pbc(s);
    }
    finish(s);
}

int md_write_grid_data_to_arrays(double* masses, struct Vector3D* positions, struct Vector3D* velocities, struct Vector3D* forces) {
    Sim& s = sim();
    const int n = std::min<int>(s.n, (int)s.h_mass.size());
    for (int i = 0; i < n; ++i) {
        masses[i] = s.h_mass[i];
        positions[i] = {s.h_pos[3 * i], s.h_pos[3 * i + 1], s.h_pos[3 * i + 2]};
        velocities[i] = {s.h_vel[3 * i], s.h_vel[3 * i + 1], s.h_vel[3 * i + 2]};
        forces[i] = {s.h_force[3 * i], s.h_force[3 * i + 1], s.h_force[3 * i + 2]};
    }
    return s.n;
}

int md_write_grid_ghost_data_to_arrays(double* masses, struct Vector3D* positions, struct Vector3D* velocities,
                                       struct Vector3D* forces) {
    Sim& s = sim();
    const int nall = std::min<int>(s.n + s.ng, (int)s.h_mass.size());
    for (int i = s.n; i < nall; ++i) {
        const int g = i - s.n;
        masses[g] = s.h_mass[i];
        positions[g] = {s.h_pos[3 * i], s.h_pos[3 * i + 1], s.h_pos[3 * i + 2]};
        velocities[g] = {s.h_vel[3 * i], s.h_vel[3 * i + 1], s.h_vel[3 * i + 2]};
        forces[g] = {s.h_force[3 * i], s.h_force[3 * i + 1], s.h_force[3 * i + 2]};
    }
    return s.n;  // the reference returns nparticles here (core/grid.impala:517)
}

int md_write_grid_aabb_data_to_arrays(double* masses, struct Vector3D* positions, struct Vector3D* velocities,
                                      struct Vector3D* forces) {
    Sim& s = sim();
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                const int idx = k * 4 + j * 2 + i;
                masses[idx] = 0.0;
                positions[idx] = {i == 0 ? s.g.aabb[0] : s.g.aabb[1], j == 0 ? s.g.aabb[2] : s.g.aabb[3],
                                  k == 0 ? s.g.aabb[4] : s.g.aabb[5]};
                velocities[idx] = {0.0, 0.0, 0.0};
                forces[idx] = {0.0, 0.0, 0.0};
            }
    return s.n;
}

// ---------------------------------------------------------------------------
// State I/O beyond the reference's array exports: the reference's legacy VTK
// particle file (vtk.h:6-91 through main.cpp:67-98) written by the library,
// and a binary restart file.

static bool vtk_file(const char* path, const std::vector<double>& m, const std::vector<Vector3D>& x,
                     const std::vector<Vector3D>& v, const std::vector<Vector3D>& f) {
    FILE* fp = fopen(path, "w");
    if (!fp) return false;
    const size_t n = x.size();
    // the layout and number format of vtk.h (std::fixed: six decimals)
    fprintf(fp, "# vtk DataFile Version 2.0\nParticle data\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    fprintf(fp, "POINTS %zu double\n", n);
    for (const Vector3D& p : x) fprintf(fp, "%f %f %f\n", p.x, p.y, p.z);
    fprintf(fp, "\n\nCELLS %zu %zu\n", n, 2 * n);
    for (size_t i = 0; i < n; ++i) fprintf(fp, "1 %zu\n", i);
    fprintf(fp, "\n\nCELL_TYPES %zu\n", n);
    for (size_t i = 0; i < n; ++i) fprintf(fp, "1\n");
    fprintf(fp, "\n\nPOINT_DATA %zu\n", m.size());
    fprintf(fp, "SCALARS mass double\nLOOKUP_TABLE default\n");
    for (double q : m) fprintf(fp, "%f\n", q);
    fprintf(fp, "\n");
    const char* names[6] = {"velocity_x", "velocity_y", "velocity_z", "force_x", "force_y", "force_z"};
    for (int c = 0; c < 6; ++c) {
        const std::vector<Vector3D>& a = c < 3 ? v : f;
        fprintf(fp, "SCALARS %s double\nLOOKUP_TABLE default\n", names[c]);
        for (const Vector3D& q : a) fprintf(fp, "%f\n", c % 3 == 0 ? q.x : (c % 3 == 1 ? q.y : q.z));
        fprintf(fp, "\n");
    }
    return fclose(fp) == 0;
}

int tmd_write_vtk(const char* path, int what) {
    Sim& s = sim();
    if (has_error()) return 0;
    if (what < 0 || what > 2) {
        set_error(__FILE__, __LINE__, "tmd_write_vtk: what must be 0 (locals), 1 (ghosts) or 2 (grid AABB)");
        return 0;
    }
    const int n = what == 0 ? s.n : (what == 1 ? s.ng : 8);
    std::vector<double> m(n);
    std::vector<Vector3D> x(n), v(n), f(n);
    if (what == 0) md_write_grid_data_to_arrays(m.data(), x.data(), v.data(), f.data());
    else if (what == 1) md_write_grid_ghost_data_to_arrays(m.data(), x.data(), v.data(), f.data());
    else md_write_grid_aabb_data_to_arrays(m.data(), x.data(), v.data(), f.data());
    if (!vtk_file(path, m, x, v, f)) {
        set_error(__FILE__, __LINE__, "tmd_write_vtk: cannot write the file");
        return 0;
    }
    return n;
}

// Restart file: a fixed header, then the locals' mass, position, velocity
// and force in FP64 (the arrays of md_initialize_grid plus the force that
// the next step's first kick uses; the (x, v) pair is what the adjoint
// checkpoints hold), then an FNV-1a checksum of everything before it.
// Ghosts, cells and lists are derived state: a restored run rebuilds them,
// as a run started from md_initialize_grid does.
struct RestartHeader {
    char magic[8];  // "TMDRST1" and a terminating zero
    int version, n, nb_cap, precision, half_mode, ghost_mode, lj_shift, pad;
    double world[6], rank[6], spacing, dem_radius;
};

static unsigned long long fnv1a(const void* p, size_t len, unsigned long long h) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < len; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

int tmd_save_state(const char* path) {
    Sim& s = sim();
    if (!ready(s)) return 0;
    if (s.pos_stale || s.ghost_stale || s.ghost_pos_stale) sync_pos_from_mirror(s);
    download(s);
    if (has_error()) return 0;
    RestartHeader h{};
    std::memcpy(h.magic, "TMDRST1", 8);
    h.version = 1;
    h.n = s.n;
    h.nb_cap = s.nb_cap;
    h.precision = s.precision;
    h.half_mode = s.half_mode;
    h.ghost_mode = s.ghost_mode;
    h.lj_shift = s.lj_shift;
    for (int d = 0; d < 6; ++d) {
        h.world[d] = s.g.world[d];
        h.rank[d] = s.g.rank[d];
    }
    h.spacing = s.g.spacing;
    h.dem_radius = s.dem_radius;
    FILE* fp = fopen(path, "wb");
    if (!fp) {
        set_error(__FILE__, __LINE__, "tmd_save_state: cannot open the file");
        return 0;
    }
    unsigned long long sum = 1469598103934665603ull;
    const size_t n = (size_t)s.n;
    bool ok = fwrite(&h, sizeof h, 1, fp) == 1;
    sum = fnv1a(&h, sizeof h, sum);
    const void* parts[4] = {s.h_mass.data(), s.h_pos.data(), s.h_vel.data(), s.h_force.data()};
    const size_t lens[4] = {n, 3 * n, 3 * n, 3 * n};
    for (int k = 0; k < 4 && ok; ++k) {
        ok = fwrite(parts[k], sizeof(double), lens[k], fp) == lens[k];
        sum = fnv1a(parts[k], sizeof(double) * lens[k], sum);
    }
    ok = ok && fwrite(&sum, sizeof sum, 1, fp) == 1;
    ok = (fclose(fp) == 0) && ok;
    if (!ok) {
        set_error(__FILE__, __LINE__, "tmd_save_state: write failed");
        return 0;
    }
    return s.n;
}

int tmd_load_state(const char* path) {
    Sim& s = sim();
    if (has_error()) return 0;
    FILE* fp = fopen(path, "rb");
    if (!fp) {
        set_error(__FILE__, __LINE__, "tmd_load_state: cannot open the file");
        return 0;
    }
    RestartHeader h{};
    bool ok = fread(&h, sizeof h, 1, fp) == 1 && std::memcmp(h.magic, "TMDRST1", 8) == 0 && h.version == 1 &&
              h.n >= 0 && h.n < (1 << 30);
    std::vector<double> m, x, v, f;
    unsigned long long sum = 1469598103934665603ull, stored = 0;
    if (ok) {
        const size_t n = (size_t)h.n;
        m.resize(n);
        x.resize(3 * n);
        v.resize(3 * n);
        f.resize(3 * n);
        sum = fnv1a(&h, sizeof h, sum);
        std::vector<double>* parts[4] = {&m, &x, &v, &f};
        for (int k = 0; k < 4 && ok; ++k) {
            ok = fread(parts[k]->data(), sizeof(double), parts[k]->size(), fp) == parts[k]->size();
            sum = fnv1a(parts[k]->data(), sizeof(double) * parts[k]->size(), sum);
        }
        ok = ok && fread(&stored, sizeof stored, 1, fp) == 1 && stored == sum;
    }
    fclose(fp);
    if (!ok) {
        set_error(__FILE__, __LINE__, "tmd_load_state: not a restart file of this version, truncated or corrupted");
        return 0;
    }
    md_initialize_grid(m.data(), reinterpret_cast<const Vector3D*>(x.data()), reinterpret_cast<const Vector3D*>(v.data()), h.n,
                       h.world, h.rank, h.spacing, 0, h.nb_cap);
    if (has_error()) return 0;
    s.h_force = f;
    s.precision = h.precision;
    s.half_mode = h.half_mode;
    s.ghost_mode = h.ghost_mode;
    s.lj_shift = h.lj_shift;
    s.dem_radius = h.dem_radius;
    upload(s);
    return has_error() ? 0 : h.n;
}

void md_distribute_particles() {
    Sim& s = sim();
    if (!ready(s)) return;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_BIN, s.stream, &ev);
    bin_particles(s);
    s.timer.end(K_BIN, s.stream, ev);
    s.nb_valid = 0;
    finish(s);
}

void md_assemble_neighborlists(bool half, double cutoff) {
    Sim& s = sim();
    if (!ready(s)) return;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_NEIGHBOR, s.stream, &ev);
    build_neighbor_lists(s, half ? 1 : 0, cutoff);
    s.timer.end(K_NEIGHBOR, s.stream, ev);
    finish(s);
}

void md_deallocate_grid() {
    if (!g_sim) return;
    Sim& s = *g_sim;
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    s.timer.collect();
    DevBuf* bufs[] = {&s.pos, &s.vel, &s.force, &s.mass, &s.sort_buf, &s.g_owner, &s.g_src, &s.g_code, &s.cell_of, &s.cell_slot,
                      &s.cell_count, &s.cell_start, &s.cell_parts, &s.cell_inv, &s.work_start, &s.work, &s.work_k, &s.work_c, &s.g_ko, &s.g_kg, &s.gh_start, &s.gh_k, &s.gh_c, &s.gh_aux, &s.tile_start, &s.spos, &s.spos2, &s.face, &s.svel, &s.svel2,
                      &s.nb_count, &s.nb_list, &s.nb_count2, &s.nb_list2,
                      &s.scan_blocks, &s.scan_aux, &s.red, &s.gbuf, &s.adj_x0, &s.adj_v0, &s.adj_ckx,
                      &s.adj_ckv, &s.adj_segx, &s.adj_segv, &s.adj_xbar, &s.adj_vbar, &s.adj_fbar, &s.adj_part,
                      &s.adj_acc};
    for (DevBuf* b : bufs) b->release();
    if (s.h_winstats) {
        (void)hipHostFree(s.h_winstats);
        s.h_winstats = nullptr;
    }
    s.windows_pending = 0;
    s.defer_window_sync = 0;
    s.max_window = 0;
    s.tile_auto = 0;
    s.stage_bytes = kPosStageBytes;
    s.stage_budget = kLdsBudget;
    s.n = s.ng = 0;
    s.cap_all = 0;
    s.nb_valid = 0;
    s.windows_valid = 0;
    s.inv_valid = 0;
    mirror_invalidate(s);
    s.pos_stale = 0;
    s.vel_mirrored = 0;
    s.vel_mirror_wanted = 0;
    s.ghost_pos_stale = 0;
    s.gsrc_valid = 0;
    s.nbinned = 0;
    s.dev_valid = false;
}

void md_print_grid() {
    Sim& s = sim();
    for (int i = 0; i < s.n && 3 * i + 2 < (int)s.h_pos.size(); ++i)
        printf("0> pos_local: %g, %g, %g\n", s.h_pos[3 * i], s.h_pos[3 * i + 1], s.h_pos[3 * i + 2]);
    for (int i = 0; i < s.n && 3 * i + 2 < (int)s.h_vel.size(); ++i)
        printf("0> vel_local: %g, %g, %g\n", s.h_vel[3 * i], s.h_vel[3 * i + 1], s.h_vel[3 * i + 2]);
}

void md_print_ghost() {
    Sim& s = sim();
    for (int i = s.n; i < s.n + s.ng && 3 * i + 2 < (int)s.h_pos.size(); ++i)
        printf("0> pos_ghost: %g, %g, %g\n", s.h_pos[3 * i], s.h_pos[3 * i + 1], s.h_pos[3 * i + 2]);
    for (int i = s.n; i < s.n + s.ng && 3 * i + 2 < (int)s.h_vel.size(); ++i)
        printf("0> vel_ghost: %g, %g, %g\n", s.h_vel[3 * i], s.h_vel[3 * i + 1], s.h_vel[3 * i + 2]);
}

void md_create_velocity(double t_request) {
    // core/runtime.impala:158-186 on the host arrays (before the upload)
    Sim& s = sim();
    if (has_error() || s.n <= 1) return;
    double vt[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < s.n; ++i)
        for (int e = 0; e < 3; ++e) vt[e] = vt[e] + s.h_vel[3 * i + e];
    for (int e = 0; e < 3; ++e) vt[e] = vt[e] / (double)s.n;
    for (int i = 0; i < s.n; ++i)
        for (int e = 0; e < 3; ++e) s.h_vel[3 * i + e] = s.h_vel[3 * i + e] - vt[e];
    const double t = host_temperature(s);
    const double factor = std::sqrt(t_request / t);
    for (int i = 0; i < s.n; ++i)
        for (int e = 0; e < 3; ++e) s.h_vel[3 * i + e] = s.h_vel[3 * i + e] * factor;
    s.dev_valid = false;
}

void md_compute_lennard_jones(bool half, double cutoff, double epsilon, double sigma) {
    Sim& s = sim();
    if (!ready(s)) return;
    compute_lj(s, half ? 1 : 0, cutoff, epsilon, sigma);
    finish(s);
}

void md_compute_dem(bool half, double cutoff, double damping_n, double damping_t, double stiffness, double friction) {
    Sim& s = sim();
    if (!ready(s)) return;
    compute_dem(s, half ? 1 : 0, cutoff, damping_n, damping_t, stiffness, friction);
    finish(s);
}

void md_mpi_initialize() {}
void md_mpi_finalize() {}

void md_synchronize_ghost_layer() {
    Sim& s = sim();
    if (!ready(s)) return;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_GHOST_SYNC, s.stream, &ev);
    refresh_ghosts(s);
    s.timer.end(K_GHOST_SYNC, s.stream, ev);
    finish(s);
}

void md_borders() {
    Sim& s = sim();
    if (!ready(s)) return;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_BORDERS, s.stream, &ev);
    build_ghosts(s);
    s.timer.end(K_BORDERS, s.stream, ev);
    s.nb_valid = 0;
    s.windows_valid = 0;
    finish(s);
}

void md_exchange_particles() {
    Sim& s = sim();
    s.ng = 0;
    s.nb_valid = 0;
    s.windows_valid = 0;
}

double md_get_send_buffer_value(int) { return 0.0; }
void md_set_recv_buffer_value(int, double) {}
void md_resize_recv_buffer_capacity(int) {}
int md_serialize_particles(double const*) { return 0; }
void md_deserialize_particles(int) {}
int md_get_world_size() { return 1; }
int md_get_world_rank() { return 0; }
int md_get_number_of_particles() { return sim().n; }
int md_get_number_of_ghost_particles() { return sim().ng; }

void md_get_node_bounding_box(double const* aabb, double (*rank)[6]) {
    for (int d = 0; d < 6; ++d) (*rank)[d] = aabb[d];
}

void md_compute_boundary_weights(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax,
                                 unsigned long int* computational_weight, unsigned long int* communication_weight) {
    // core/runtime.impala:72-95: atoms and ghosts inside the AABB (inclusive
    // bounds, core/grid.impala:413-417), counted on the host mirror
    Sim& s = sim();
    unsigned long int c = 0, g = 0;
    const int nall = std::min<int>(s.n + s.ng, (int)(s.h_pos.size() / 3));
    for (int i = 0; i < nall; ++i) {
        const double* p = &s.h_pos[3 * i];
        const bool in = p[0] >= xmin && p[0] <= xmax && p[1] >= ymin && p[1] <= ymax && p[2] >= zmin && p[2] <= zmax;
        if (in) (i < s.n ? c : g)++;
    }
    *computational_weight = c;
    *communication_weight = g;
}

void md_update_neighborhood(int, int, int const*, int const*, double const*) {}

// ForceIters = list entries visited by the last pair-force evaluation,
// RsqIters = those inside its cutoff, IntegrationIters = atoms of the last
// integration (core/compute.impala:1-48, 97-116; scalar counts, one per
// entry and atom).  Both list counts are taken from the current list and
// positions, which are those of the last force evaluation: the final
// integration moves only velocities.
void md_report_iterations() {
    Sim& s = sim();
    long long entries = 0, in_cut = 0;
    if (s.nb_valid && s.stream && !has_error()) {
        std::vector<int> c(std::max(s.n, 1));
        TMD_HIP(hipMemcpyAsync(c.data(), s.nb_count.p, sizeof(int) * s.n, hipMemcpyDeviceToHost, s.stream));
        TMD_HIP(hipStreamSynchronize(s.stream));
        for (int i = 0; i < s.n; ++i) entries += c[i];
        if (s.last_force_rc > 0.0) {
            if (s.pos_stale || s.ghost_stale) sync_pos_from_mirror(s);
            double pe = 0.0, vir = 0.0;
            lj_energy_virial(s, s.last_force_rc, 1.0, 1.0, &pe, &vir, &in_cut);
        }
    }
    s.force_iters = entries;
    s.rsq_iters = in_cut;
    printf("ForceIters RsqIters IntegrationIters\n");
    printf("%lld %lld %lld \n", entries, in_cut, s.integration_iters);
    fflush(stdout);
}

void md_report_particles() {
    Sim& s = sim();
    printf("# Local # Ghost Temp\n");
    printf("%d %d %g \n", s.n, s.ng, (int)s.h_vel.size() >= 3 * s.n && s.n > 1 ? host_temperature(s) : 0.0);
    fflush(stdout);
}

void md_report_time(double total, double force, double neigh, double comm, double balance, double other) {
    printf("Total Force Neigh Comm Balance Other\n");
    printf("%g %g %g %g %g %g \n", total, force, neigh, comm, balance, other);
    fflush(stdout);
}

void md_barrier() {
    Sim& s = sim();
    if (s.stream && !has_error()) TMD_HIP(hipStreamSynchronize(s.stream));
}

// ===========================================================================
// extensions
const char* tmd_last_error() { return g_has_error ? g_error.c_str() : nullptr; }

void tmd_clear_error() {
    g_has_error = false;
    g_error.clear();
}

int tmd_device_info(char* buf, int len) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        if (buf && len > 0) buf[0] = 0;
        return 0;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    // some runtimes report no marketing name: the architecture alone then
    if (buf && len > 0) {
        if (prop.name[0]) snprintf(buf, len, "%s (%s)", prop.name, prop.gcnArchName);
        else snprintf(buf, len, "%s", prop.gcnArchName);
    }
    return prop.multiProcessorCount;
}

void tmd_set_precision(int mode) { sim().precision = mode == 2 ? PREC_FP32_KAHAN : (mode ? PREC_FP32_MIXED : PREC_FP64); }
void tmd_set_half_mode(int mode) { sim().half_mode = mode ? HALF_REF_LITERAL : HALF_GHOST_FULL; }
void tmd_set_ghost_mode(int mode) { sim().ghost_mode = mode ? GHOST_REF_LAG : GHOST_FRESH; }
void tmd_set_sync(int on) { sim().sync_calls = on ? 1 : 0; }
void tmd_set_dem_radius(double r) { sim().dem_radius = r; }
void tmd_set_lj_shift(int on) { sim().lj_shift = on ? 1 : 0; }
void tmd_set_sort_particles(int on) { sim().sort_particles = on ? 1 : 0; }
void tmd_set_adjoint_precision(int mode) { sim().adj_precision = mode < 0 ? -1 : (mode ? PREC_FP32_MIXED : PREC_FP64); }

void tmd_set_tile_threads(int threads) {
    Sim& s = sim();
    if (threads != 0 && (threads < 64 || threads > kMaxTileThreads || threads % 64 != 0)) {
        set_error(__FILE__, __LINE__, "tmd_set_tile_threads: threads must be 0 (automatic) or a multiple of 64 in [64, 1024]");
        return;
    }
    s.tile_threads_req = threads;
    if (threads > 0) s.tile_threads = threads;
}

void tmd_set_tile(int tx, int ty, int tz) {
    Sim& s = sim();
    s.tile_req[0] = tx;
    s.tile_req[1] = ty;
    s.tile_req[2] = tz;
    s.tile_auto = 0;
    s.windows_valid = 0;
    s.nb_valid = 0;
}

static void thermo(double rc, double eps, double sigma, double* out, bool eshift) {
    Sim& s = sim();
    for (int k = 0; k < 6; ++k) out[k] = 0.0;
    if (!ready(s)) return;
    hipEvent_t ev = nullptr;
    s.timer.begin(K_THERMO, s.stream, &ev);
    const double sum_mv2 = kinetic_sum(s);
    double pe = 0.0, vir = 0.0;
    long long pairs = 0;
    lj_energy_virial(s, rc, eps, sigma, &pe, &vir, &pairs, eshift);
    s.timer.end(K_THERMO, s.stream, ev);
    const double T = sum_mv2 * (1.0 / (double)(s.n * 3 - 3));
    const double vol = s.g.len[0] * s.g.len[1] * s.g.len[2];
    s.pe = pe;
    s.virial = vir;
    s.pairs = pairs;
    out[0] = T;
    out[1] = 0.5 * sum_mv2;
    out[2] = pe;
    out[3] = vir;
    out[4] = (sum_mv2 + vir) / (3.0 * vol);
    out[5] = (double)pairs;
}

void tmd_thermo(double rc, double eps, double sigma, double* out) { thermo(rc, eps, sigma, out, false); }

void tmd_thermo_eshift(double rc, double eps, double sigma, double* out) { thermo(rc, eps, sigma, out, true); }

void tmd_thermo_noalloc(double rc, double eps, double sigma, double* out) {
    Sim& s = sim();
    for (int k = 0; k < 6; ++k) out[k] = 0.0;
    if (has_error()) return;
    if (!s.stream || !s.dev_valid) {
        set_error(__FILE__, __LINE__, "allocation-free thermo: the state is not on the device "
                                      "(md_copy_data_to_accelerator first)");
        return;
    }
    const int ntiles = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    const size_t need = std::max(3 * (size_t)std::max(ntiles, 0) + 8, (size_t)div_up(std::max(s.n, 1), 256) + 8);
    if (need > kScratchDoubles) {
        char msg[200];
        snprintf(msg, sizeof msg, "allocation-free thermo: %d atoms in %d tiles need %zu doubles of scratch, more "
                                  "than the %zu it holds", s.n, ntiles, need, kScratchDoubles);
        set_error(__FILE__, __LINE__, msg);
        return;
    }
    double* buf = scratch_ptr();
    if (!buf) return;
    AllocFree guard;
    // the same reductions as tmd_thermo, one after the other in the scratch
    const double sum_mv2 = kinetic_sum(s, buf);
    double pe = 0.0, vir = 0.0;
    long long pairs = 0;
    lj_energy_virial(s, rc, eps, sigma, &pe, &vir, &pairs, false, buf);
    if (has_error()) return;
    const double vol = s.g.len[0] * s.g.len[1] * s.g.len[2];
    s.pe = pe;
    s.virial = vir;
    s.pairs = pairs;
    out[0] = sum_mv2 * (1.0 / (double)(s.n * 3 - 3));
    out[1] = 0.5 * sum_mv2;
    out[2] = pe;
    out[3] = vir;
    out[4] = (sum_mv2 + vir) / (3.0 * vol);
    out[5] = (double)pairs;
}

void tmd_get_grid_info(int* out) {
    Sim& s = sim();
    out[0] = s.g.nx; out[1] = s.g.ny; out[2] = s.g.nz; out[3] = s.g.ncells;
    out[4] = s.n; out[5] = s.ng; out[6] = s.nb_cap;
}

void tmd_get_tile_info(int* out) {
    const Sim& s = sim();
    out[0] = s.tiles.tx; out[1] = s.tiles.ty; out[2] = s.tiles.tz;
    out[3] = s.tiles.ntx * s.tiles.nty * s.tiles.ntz;
    out[4] = s.tile_threads; out[5] = s.max_window; out[6] = s.stage_bytes;
}

int tmd_get_ghost_table(int* owner, int* image3, int* source, int* flag3) {
    Sim& s = sim();
    if (!s.stream || has_error() || s.ng == 0) return s.ng;
    std::vector<int> o(s.ng), src(s.ng), code(s.ng);
    TMD_HIP(hipMemcpy(o.data(), s.g_owner.p, sizeof(int) * s.ng, hipMemcpyDeviceToHost));
    TMD_HIP(hipMemcpy(src.data(), s.g_src.p, sizeof(int) * s.ng, hipMemcpyDeviceToHost));
    TMD_HIP(hipMemcpy(code.data(), s.g_code.p, sizeof(int) * s.ng, hipMemcpyDeviceToHost));
    for (int g = 0; g < s.ng; ++g) {
        if (owner) owner[g] = o[g];
        if (source) source[g] = src[g];
        const int ic = code[g] & 0xff, fc = (code[g] >> 8) & 0xff;
        if (image3) { image3[3 * g] = ic % 3 - 1; image3[3 * g + 1] = (ic / 3) % 3 - 1; image3[3 * g + 2] = ic / 9 - 1; }
        if (flag3) { flag3[3 * g] = fc % 3 - 1; flag3[3 * g + 1] = (fc / 3) % 3 - 1; flag3[3 * g + 2] = fc / 9 - 1; }
    }
    return s.ng;
}

int tmd_get_cells(int* start, int* atoms, int* atom_cell) {
    Sim& s = sim();
    if (!s.stream || has_error()) return 0;
    const int nc = s.g.ncells;
    std::vector<int> st(nc + 1);
    TMD_HIP(hipMemcpy(st.data(), s.cell_start.p, sizeof(int) * (nc + 1), hipMemcpyDeviceToHost));
    if (start) std::copy(st.begin(), st.end(), start);
    if (atoms && st[nc] > 0) TMD_HIP(hipMemcpy(atoms, s.cell_parts.p, sizeof(int) * st[nc], hipMemcpyDeviceToHost));
    if (atom_cell && s.n + s.ng > 0) TMD_HIP(hipMemcpy(atom_cell, s.cell_of.p, sizeof(int) * (s.n + s.ng), hipMemcpyDeviceToHost));
    return st[nc];
}

int tmd_get_neighbor_lists(int* counts, int* list, int stride) {
    // lists hold 16-bit offsets into each tile's window; resolve them to atom
    // indices with the same window construction as the kernels
    Sim& s = sim();
    if (!s.stream || has_error() || !s.nb_valid) return 0;
    const int n = s.n, cap = s.nb_cap, ntile64 = (n + 63) / 64, nc = s.g.ncells;
    const TileGeom g = s.tiles;
    const int ntiles = g.ntx * g.nty * g.ntz;
    std::vector<int> work(std::max(n, 1)), cnt(std::max(n, 1)), st(nc + 1), ts(ntiles + 1);
    TMD_HIP(hipMemcpy(work.data(), s.work.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    TMD_HIP(hipMemcpy(cnt.data(), s.nb_count.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    TMD_HIP(hipMemcpy(st.data(), s.cell_start.p, sizeof(int) * (nc + 1), hipMemcpyDeviceToHost));
    TMD_HIP(hipMemcpy(ts.data(), s.tile_start.p, sizeof(int) * (ntiles + 1), hipMemcpyDeviceToHost));
    std::vector<int> parts(std::max(st[nc], 1));
    TMD_HIP(hipMemcpy(parts.data(), s.cell_parts.p, sizeof(int) * st[nc], hipMemcpyDeviceToHost));
    std::vector<unsigned short> lst(list ? (size_t)ntile64 * 64 * cap : 1);
    if (list) TMD_HIP(hipMemcpy(lst.data(), s.nb_list.p, sizeof(unsigned short) * lst.size(), hipMemcpyDeviceToHost));
    if (has_error()) return 0;
    std::vector<int> seg_beg, seg_off;
    for (int tile = 0; tile < ntiles; ++tile) {
        const int w0 = ts[tile], w1 = ts[tile + 1];
        if (w0 == w1) continue;
        const int tx_i = tile % g.ntx, ty_i = (tile / g.ntx) % g.nty, tz_i = tile / (g.ntx * g.nty);
        const int cx0 = 1 + tx_i * g.tx, cy0 = 1 + ty_i * g.ty, cz0 = 1 + tz_i * g.tz;
        const int wy = g.ty + 2, nseg = wy * (g.tz + 2);
        seg_beg.assign(nseg, 0);
        seg_off.assign(nseg + 1, 0);
        for (int q = 0; q < nseg; ++q) {
            const int y = cy0 - 1 + q % wy, z = cz0 - 1 + q / wy;
            const int xlo = cx0 - 1, xhi = std::min(cx0 + g.tx, g.nx - 1);
            int b = 0, len = 0;
            if (y < g.ny && z < g.nz) {
                const int row = (z * g.ny + y) * g.nx;
                b = st[row + xlo];
                len = st[row + xhi + 1] - b;
            }
            seg_beg[q] = b;
            seg_off[q + 1] = seg_off[q] + len;
        }
        for (int w = w0; w < w1; ++w) {
            const int i = work[w];
            counts[i] = cnt[w];
            if (!list) continue;
            const int m = std::min(std::min(cnt[w], stride), cap);
            for (int k = 0; k < m; ++k) {
                const int t = lst[(size_t)(w >> 6) * cap * 64 + (size_t)(k >> 3) * 512 + (size_t)(w & 63) * 8 + (size_t)(k & 7)];
                int q = 0;
                while (q + 1 < nseg && seg_off[q + 1] <= t) ++q;
                list[(size_t)i * stride + k] = parts[seg_beg[q] + (t - seg_off[q])];
            }
        }
    }
    return n;
}

static int run_loop(int steps, int reneigh_every, double dt, bool half, double cutoff, double skin, int dem, double p0, double p1,
                    double p2, double p3) {
    Sim& s = sim();
    if (!ready(s)) return 0;
    if (reneigh_every <= 0) {
        set_error(__FILE__, __LINE__, "run: reneighbouring interval must be positive");
        return 0;
    }
    const int hf = half ? 1 : 0;
    const int saved_sync = s.sync_calls;
    s.sync_calls = 0;
    if (!s.nb_valid || s.nb_half != hf) rebuild(s, hf, cutoff + skin);
    int done = 0;
    if (!dem && s.ghost_mode == GHOST_FRESH) {
        // LJ: the force pass also performs the final kick of its step and
        // the first kick + drift of the next one, so a step is a ghost/mirror
        // refresh (or a rebuild) and one force pass.  Automatic tiles are
        // sized for that pass's positions-only window slots.
        s.stage_bytes = kPosStageBytes;
        s.stage_budget = kLdsBudget;
        if (steps > 0) initial_integration(s, dt);
        mirror_velocities(s, true);
        s.vel_mirror_wanted = 1;  // the rebuild's mirror pass moves them again
        for (int j = 0; j < steps && !has_error(); ++j) {
            if ((j + 1) % reneigh_every == 0) {
                rebuild(s, hf, cutoff + skin);  // its ghost selection pass restores pos and vel from the mirror
            } else {
                hipEvent_t ev = nullptr;
                s.timer.begin(K_GHOST_SYNC, s.stream, &ev);
                step_refresh(s);
                s.timer.end(K_GHOST_SYNC, s.stream, ev);
            }
            compute_lj(s, hf, cutoff, p0, p1, dt, j + 1 < steps);
            ++done;
        }
        s.vel_mirror_wanted = 0;
        sync_pos_from_mirror(s);
    } else if (dem && s.ghost_mode == GHOST_FRESH) {
        // DEM: positions and velocities (ghosts included) live in the cell-
        // ordered mirrors; the force pass ends each step and runs the next
        // step's kick and drift (compute_dem_fused), and a step between
        // rebuilds is the mirror-only ghost pass and one force pass.  That
        // pass stages positions and velocities (48 B per window slot), so
        // automatic tiles are sized for three of its blocks per CU from the
        // next rebuild on.
        s.stage_bytes = kDemStageBytes;
        s.stage_budget = kDemLdsShare;
        retune_tile_edge(s);
        if (steps > 0) initial_integration(s, dt);
        for (int j = 0; j < steps && !has_error(); ++j) {
            hipEvent_t ev = nullptr;
            if ((j + 1) % reneigh_every == 0) {
                // the rebuild's ghost selection pass copies the locals'
                // positions and velocities back from the mirrors (as in the
                // LJ loop); its ghosts are made anew from them
                if (s.dem_mirrored) {
                    s.pos_stale = 1;
                    s.vel_mirrored = 1;
                    s.dem_mirrored = 0;
                }
                rebuild(s, hf, cutoff + skin);
            } else if (s.dem_mirrored) {
                // the previous step pass wrote the ghosts too, unless it
                // could not (no ghost index)
                if (!s.dem_ghosts_done) {
                    s.timer.begin(K_GHOST_SYNC, s.stream, &ev);
                    dem_ghost_step(s);
                    s.timer.end(K_GHOST_SYNC, s.stream, ev);
                }
            } else {
                s.timer.begin(K_GHOST_SYNC, s.stream, &ev);
                refresh_ghosts(s);
                s.timer.end(K_GHOST_SYNC, s.stream, ev);
            }
            if (!s.dem_mirrored) {
                s.timer.begin(K_MIRROR, s.stream, &ev);
                dem_mirror(s);
                s.timer.end(K_MIRROR, s.stream, ev);
            }
            compute_dem_fused(s, hf, cutoff, p0, p1, p2, p3, dt, j + 1 < steps);
            ++done;
        }
        dem_unmirror(s);  // only after an error: the last step returns to atom order itself
    } else if (dem) {
        // DEM with the reference's lagging ghost refresh: atom order, the
        // step's final kick and the next step's kick and drift one pass
        // after the force
        if (steps > 0) initial_integration(s, dt);
        for (int j = 0; j < steps && !has_error(); ++j) {
            if ((j + 1) % reneigh_every == 0) {
                rebuild(s, hf, cutoff + skin);
            } else {
                hipEvent_t ev = nullptr;
                s.timer.begin(K_GHOST_SYNC, s.stream, &ev);
                refresh_ghosts(s);
                s.timer.end(K_GHOST_SYNC, s.stream, ev);
            }
            compute_dem(s, hf, cutoff, p0, p1, p2, p3, dt, j + 1 < steps);
            ++done;
        }
    } else {
        for (int j = 0; j < steps && !has_error(); ++j) {
            initial_integration(s, dt);
            if ((j + 1) % reneigh_every == 0) {
                rebuild(s, hf, cutoff + skin);
            } else {
                hipEvent_t ev = nullptr;
                s.timer.begin(K_GHOST_SYNC, s.stream, &ev);
                refresh_ghosts(s);
                s.timer.end(K_GHOST_SYNC, s.stream, ev);
            }
            compute_lj(s, hf, cutoff, p0, p1, dt);  // final kick fused
            ++done;
        }
    }
    s.sync_calls = saved_sync;
    if (!has_error()) TMD_HIP(hipStreamSynchronize(s.stream));
    s.integration_iters = s.n;
    return has_error() ? 0 : done;
}

int tmd_run_lj(int steps, int reneigh_every, double dt, bool half, double cutoff, double skin, double eps, double sigma) {
    return run_loop(steps, reneigh_every, dt, half, cutoff, skin, 0, eps, sigma, 0.0, 0.0);
}

int tmd_run_dem(int steps, int reneigh_every, double dt, bool half, double cutoff, double skin, double damping_n, double damping_t,
                double stiffness, double friction) {
    return run_loop(steps, reneigh_every, dt, half, cutoff, skin, 1, damping_n, damping_t, stiffness, friction);
}

int tmd_traj_loss(int steps, int reneigh_every, double dt, double cutoff, double skin, double epsilon, double sigma,
                  double* out2) {
    Sim& s = sim();
    out2[0] = out2[1] = 0.0;
    if (!ready(s)) return 0;
    const int saved_sync = s.sync_calls;
    s.sync_calls = 0;
    const bool ok = traj_loss(s, steps, reneigh_every, dt, cutoff, skin, epsilon, sigma, &out2[0], &out2[1]);
    s.sync_calls = saved_sync;
    return ok ? 1 : 0;
}

int tmd_traj_adjoint(int steps, int reneigh_every, double dt, double cutoff, double skin, double epsilon, double sigma,
                     double u_target, double p_target, double* out3, double* grad_theta2, struct Vector3D* grad_x0,
                     struct Vector3D* grad_v0) {
    Sim& s = sim();
    out3[0] = out3[1] = out3[2] = 0.0;
    grad_theta2[0] = grad_theta2[1] = 0.0;
    if (!ready(s)) return 0;
    const int saved_sync = s.sync_calls;
    s.sync_calls = 0;
    const bool ok = traj_adjoint(s, steps, reneigh_every, dt, cutoff, skin, epsilon, sigma, u_target, p_target, out3,
                                 grad_theta2, reinterpret_cast<double*>(grad_x0), reinterpret_cast<double*>(grad_v0));
    s.sync_calls = saved_sync;
    return ok ? 1 : 0;
}

void tmd_profile_enable(int on) {
    Sim& s = sim();
    if (!on) s.timer.collect();
    s.timer.enabled = on != 0;
}

void tmd_profile_get(int kernel, double* total_ms, long long* count) {
    Sim& s = sim();
    s.timer.collect();
    *total_ms = 0.0;
    *count = 0;
    if (kernel >= 0 && kernel < K_NKERNELS) {
        *total_ms = s.timer.total_ms[kernel];
        *count = s.timer.count[kernel];
    }
}

void tmd_profile_reset() { sim().timer.reset(); }

int tmd_generate_fcc(int nx, int ny, int nz, double potmin, int half, double* mass, struct Vector3D* pos, struct Vector3D* vel,
                     int cap) {
    // initialize.h:91-201 with aabb = rank_aabb = [0, n * potmin] and
    // spacing = potmin / 2 (main.cpp:515-518, 670)
    double aabb[6] = {0.0, nx * potmin, 0.0, ny * potmin, 0.0, nz * potmin};
    const double spacing = potmin * 0.5;
    int nvertices[3], nxyz[3];
    for (int d = 0; d < 3; ++d) {
        nvertices[d] = (int)((aabb[d * 2 + 1] - aabb[d * 2 + 0]) / spacing);
        nxyz[d] = nvertices[d] / 2;
    }
    int ilo = std::max((int)(aabb[0] / spacing - 1), 0), ihi = std::min((int)(aabb[1] / spacing + 1), nvertices[0] - 1);
    int jlo = std::max((int)(aabb[2] / spacing - 1), 0), jhi = std::min((int)(aabb[3] / spacing + 1), nvertices[1] - 1);
    int klo = std::max((int)(aabb[4] / spacing - 1), 0), khi = std::min((int)(aabb[5] / spacing + 1), nvertices[2] - 1);
    const int sub = 8;
    int sx = 0, sy = 0, sz = 0, ox = 0, oy = 0, oz = 0, count = 0;
    while (oz * sub <= khi) {
        const int k = oz * sub + sz, j = oy * sub + sy, i = ox * sub + sx;
        if ((i + j + k) % 2 == 0 && i >= ilo && i <= ihi && j >= jlo && j <= jhi && k >= klo && k <= khi) {
            const double px = aabb[0] + i * spacing, py = aabb[2] + j * spacing, pz = aabb[4] + k * spacing;
            const bool inside = px >= aabb[0] && px < aabb[1] - 0.00001 && py >= aabb[2] && py < aabb[3] - 0.00001 &&
                                pz >= aabb[4] && pz < aabb[5] - 0.00001;
            if (inside && (!half || py < aabb[2] + ((jhi - jlo) / 2) * spacing + 0.00001)) {
                int seed = k * (2 * nxyz[1]) * (2 * nxyz[0]) + j * (2 * nxyz[0]) + i + 1;
                double v3[3];
                for (int c = 0; c < 3; ++c) {
                    for (int r = 0; r < 5; ++r) pm_random(&seed);
                    v3[c] = pm_random(&seed);
                }
                if (count < cap) {
                    mass[count] = 1.0;
                    pos[count] = {px, py, pz};
                    vel[count] = {v3[0], v3[1], v3[2]};
                }
                ++count;
            }
        }
        sx++;
        if (sx == sub) { sx = 0; sy++; }
        if (sy == sub) { sy = 0; sz++; }
        if (sz == sub) { sz = 0; ox++; }
        if (ox * sub > ihi) { ox = 0; oy++; }
        if (oy * sub > jhi) { oy = 0; oz++; }
    }
    return count;
}

int tmd_generate_granular_gas(int nc, double potmin, unsigned long long seed, double vmax, double* mass, struct Vector3D* pos,
                              struct Vector3D* vel, int cap) {
    const double spacing = potmin * 0.5;
    const int nv = 2 * nc;
    int count = 0;
    for (int k = 0; k < nv; ++k)
        for (int j = 0; j < nv; ++j)
            for (int i = 0; i < nv; ++i) {
                if ((i + j + k) % 2 != 0 || i + j >= nv) continue;
                if (count < cap) {
                    mass[count] = 1.0;
                    pos[count] = {i * spacing, j * spacing, k * spacing};
                    const unsigned long long site = ((unsigned long long)k * nv + j) * nv + i;
                    double v3[3];
                    for (int c = 0; c < 3; ++c) {
                        const unsigned long long h = mix64(seed ^ mix64(site * 3ull + (unsigned long long)c));
                        const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
                        v3[c] = vmax * (2.0 * u - 1.0);
                    }
                    vel[count] = {v3[0], v3[1], v3[2]};
                }
                ++count;
            }
    return count;
}

}  // extern "C"
