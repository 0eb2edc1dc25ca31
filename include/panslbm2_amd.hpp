/*
 * include/panslbm2_amd.hpp -- C++ drop-in for the reference's templated lattice API, running on
 * the MI355X through the C ABI of panslbm2_amd.h (link with -lpanslbm2_hip).
 *
 * The reference's drivers (production/heatsink.cpp, test/nsadsens.cpp, ...) call header-only
 * templates: a lattice class (src/particle/d2q9.h:24-158, d3q15.h:24-249) and free functions in
 * namespaces NS, AD, ANS, AAD (the headers under src/equation/) that take the lattice, raw host field arrays
 * and host lambdas.  This header declares the same names with the same signatures in namespace
 * PANSLBM2, so such a driver switches by changing its #includes.  Each function cites the
 * reference function it replaces.
 *
 * How the calls map onto the device:
 *  - A lattice owns its populations on the GPU.  Its public f0/f host arrays (reference layout
 *    f0[N], f[(nc-1)*idx + c-1]) are filled by Download() and pushed by Upload(); nothing else
 *    reads or writes them.
 *  - Stream/iStream, the boundary-condition calls and SmoothCorner are recorded, not run, and
 *    the next collide runs "stream + boundary calls + corner smoothing + collide" as one fused
 *    pass over memory.  Any other call first runs what is pending.  Identical recordings (the
 *    same lambdas giving the same values, the same arrays) reuse one uploaded program.
 *  - Lambdas are evaluated on the host at the call, on the domain faces, in the reference's
 *    face order (xmin, xmax, ymin, ymax[, zmin, zmax]), with global coordinates; value lambdas
 *    only where the bctype lambda selects the node.
 *  - Every host field array passed in gets a device copy (keyed by its address, sized by the
 *    call).  gpu::SetCoherence picks how the two stay in step:
 *      Coherence::Host (default): inputs are compared with what was last uploaded on every call
 *        and re-uploaded when the host changed them; outputs are downloaded when a call writes
 *        them (issave).  Host arrays always hold what the reference would leave there.
 *      Coherence::Device: the device copy is authoritative between calls.  Read a host array
 *        the device wrote only after gpu::Sync(ptr) or gpu::SyncAll(); after writing an array
 *        on the host, call gpu::Invalidate(ptr).  Arrays the device never wrote (alpha,
 *        diffusivity, ...) are re-checked at every InitialCondition call.  Residual and the
 *        sensitivities read the device copies, so a converged solve moves no field to the host.
 *    Sensitivity outputs (dfds, accumulated +=) are uploaded before and downloaded after the
 *    call in both modes.
 *  - Errors throw PANSLBM2::gpu::Error carrying lbm_last_error() (the reference asserts).
 *  - T must be double: the device path is FP64.  One GPU holds the whole domain, so the
 *    constructors accept only PEid = 0 and mx = my = mz = 1.
 */
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <list>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "panslbm2_amd.h"

namespace PANSLBM2 {

namespace gpu {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc, const char *what) {
    if (rc != LBM_OK) throw Error(std::string(what) + ": " + lbm_last_error());
}

enum class Coherence { Host, Device };

// Process-wide state: the context (device 0), the device copies of host arrays and the
// recording of stream / boundary / smoothing calls awaiting the next collide.
class Runtime {
public:
    static Runtime &get() {
        // never destroyed: the HIP runtime may already be torn down when statics are destroyed
        static Runtime *rt = new Runtime();
        return *rt;
    }
    lbm_context *ctx() const { return ctx_; }
    Coherence coherence = Coherence::Host;
    bool ans_compat = true;  // keep the reference's ANS y-edge velocity formula

    // ---- device copies of host arrays -------------------------------------------------------
    // nc > 0 marks a population snapshot: host layout nc*idx + c, device layout c*N + idx.
    const double *in(const double *h, size_t count, int nc = 0) {
        if (!h) return nullptr;
        bool fresh = false;
        Mirror &m = mirror(h, count, nc, fresh);
        if (fresh) return m.dev;
        if (m.stale) {
            flush();  // recorded boundary calls must read the values they were given
            upload(h, m);
        } else if (coherence == Coherence::Host || m.recheck) {
            refresh(h, m);
        }
        m.recheck = false;
        return m.dev;
    }
    double *out(double *h, size_t count, int nc = 0) {
        if (!h) return nullptr;
        bool fresh = false;
        return mirror(h, count, nc, fresh).dev;
    }
    double *inout(double *h, size_t count) {
        if (!h) return nullptr;
        bool fresh = false;
        Mirror &m = mirror(h, count, 0, fresh);
        if (!fresh && m.stale) {
            flush();
            upload(h, m);
        } else if (!fresh) {
            refresh(h, m);
        }
        return m.dev;
    }
    // the device has written these host arrays
    void wrote(std::initializer_list<double *> hs, bool force_download = false) {
        for (double *h : hs) {
            if (!h) continue;
            Mirror &m = mirrors_.at(h);
            m.device_owned = true;
            m.device_newer = true;
            m.stale = m.recheck = false;  // the device copy is now the newest
            if (coherence == Coherence::Host || force_download) download(h, m);
        }
    }
    void sync(const double *h) {
        auto it = mirrors_.find(h);
        if (it != mirrors_.end() && it->second.device_newer) download(const_cast<double *>(h), it->second);
    }
    void sync_all() {
        for (auto &kv : mirrors_)
            if (kv.second.device_newer) download(const_cast<double *>(kv.first), kv.second);
    }
    void invalidate(const double *h) {
        auto it = mirrors_.find(h);
        if (it == mirrors_.end()) return;
        it->second.device_newer = false;
        it->second.stale = true;
    }
    void forget(const double *h) {
        auto it = mirrors_.find(h);
        if (it == mirrors_.end()) return;
        flush();
        drop_programs_using(nullptr);  // a cached program may hold the device address
        check(lbm_device_free(ctx_, it->second.dev), "lbm_device_free");
        mirrors_.erase(it);
    }
    // start of a solve: arrays the device never wrote are compared with the host again
    void recheck_host_owned() {
        for (auto &kv : mirrors_)
            if (!kv.second.device_owned) kv.second.recheck = true;
    }

    // ---- recording ----------------------------------------------------------------------------
    // Record a stream / boundary / smoothing call.  A call the open recording cannot take (a
    // third lattice, a second stream of one lattice, ...) runs what is recorded so far and
    // starts a new recording.
    template <class Call>
    void record(lbm_lattice *L, const char *what, Call call, lbm_lattice *L2 = nullptr) {
        begin();
        int rc = call();
        if (rc == LBM_ERR_UNSUPPORTED) {
            flush();
            begin();
            rc = call();
        }
        check(rc, what);
        for (lbm_lattice *x : {L, L2})
            if (x && std::find(lats_.begin(), lats_.end(), x) == lats_.end()) lats_.push_back(x);
    }
    // the pending program for a collide (nullptr when nothing is recorded)
    lbm_program *take() {
        if (!recording_) return nullptr;
        lbm_program *P = nullptr;
        recording_ = false;
        check(lbm_record_end_deferred(ctx_, &P), "lbm_record_end_deferred");
        for (auto it = cache_.begin(); it != cache_.end(); ++it) {
            if (lbm_program_equal(it->prog, P) == 1) {
                check(lbm_program_destroy(P), "lbm_program_destroy");
                cache_.splice(cache_.begin(), cache_, it);
                return cache_.front().prog;
            }
        }
        cache_.push_front(Cached{P, lats_});
        while (cache_.size() > kCachedPrograms) {
            check(lbm_program_destroy(cache_.back().prog), "lbm_program_destroy");
            cache_.pop_back();
        }
        return P;
    }
    // run whatever is recorded on its own
    void flush() {
        if (lbm_program *P = take()) check(lbm_program_apply(P), "lbm_program_apply");
    }
    void lattice_gone(lbm_lattice *L) {
        flush();
        drop_programs_using(L);
    }

private:
    struct Mirror {
        double *dev = nullptr;
        size_t count = 0;
        int nc = 0;
        std::vector<double> shadow;  // host-layout contents at the last transfer
        bool device_owned = false;   // a call has written this array on the device
        bool device_newer = false;   // the device holds values the host has not been given
        bool recheck = false;        // compare with the host at the next use (Device mode)
        bool stale = false;          // the host wrote it: upload at the next use
    };
    struct Cached {
        lbm_program *prog;
        std::vector<lbm_lattice *> lats;
    };
    static constexpr size_t kCachedPrograms = 8;

    Runtime() { check(lbm_context_create(0, &ctx_), "lbm_context_create"); }

    void begin() {
        if (recording_) return;
        check(lbm_record_begin(ctx_), "lbm_record_begin");
        recording_ = true;
        lats_.clear();
    }
    void drop_programs_using(lbm_lattice *L) {
        for (auto it = cache_.begin(); it != cache_.end();) {
            bool uses = L == nullptr;
            for (lbm_lattice *x : it->lats) uses |= (x == L);
            if (uses) {
                check(lbm_program_destroy(it->prog), "lbm_program_destroy");
                it = cache_.erase(it);
            } else {
                ++it;
            }
        }
    }
    Mirror &mirror(const double *h, size_t count, int nc, bool &fresh) {
        auto it = mirrors_.find(h);
        if (it != mirrors_.end() && it->second.count == count && it->second.nc == nc) return it->second;
        if (it != mirrors_.end()) forget(h);  // same address, new array
        Mirror m;
        m.count = count;
        m.nc = nc;
        void *d = nullptr;
        check(lbm_device_alloc(ctx_, sizeof(double) * count, &d), "lbm_device_alloc");
        m.dev = static_cast<double *>(d);
        Mirror &ref = mirrors_.emplace(h, std::move(m)).first->second;
        upload(h, ref);
        fresh = true;
        return ref;
    }
    void refresh(const double *h, Mirror &m) {
        if (!std::memcmp(h, m.shadow.data(), sizeof(double) * m.count)) return;
        flush();  // recorded boundary calls must read the values they were given
        upload(h, m);
    }
    void upload(const double *h, Mirror &m) {
        m.shadow.assign(h, h + m.count);
        if (m.nc) {
            std::vector<double> t(m.count);
            size_t N = m.count / m.nc;
            for (size_t idx = 0; idx < N; ++idx)
                for (int c = 0; c < m.nc; ++c) t[c * N + idx] = h[m.nc * idx + c];
            check(lbm_copy_to_device(ctx_, m.dev, t.data(), sizeof(double) * m.count), "upload");
        } else {
            check(lbm_copy_to_device(ctx_, m.dev, h, sizeof(double) * m.count), "upload");
        }
        m.device_newer = false;
        m.stale = false;
    }
    void download(double *h, Mirror &m) {
        if (m.nc) {
            std::vector<double> t(m.count);
            check(lbm_copy_to_host(ctx_, t.data(), m.dev, sizeof(double) * m.count), "download");
            size_t N = m.count / m.nc;
            for (size_t idx = 0; idx < N; ++idx)
                for (int c = 0; c < m.nc; ++c) h[m.nc * idx + c] = t[c * N + idx];
        } else {
            check(lbm_copy_to_host(ctx_, h, m.dev, sizeof(double) * m.count), "download");
        }
        m.shadow.assign(h, h + m.count);
        m.device_newer = false;
        m.stale = false;
    }

    lbm_context *ctx_ = nullptr;
    std::unordered_map<const double *, Mirror> mirrors_;
    bool recording_ = false;
    std::vector<lbm_lattice *> lats_;
    std::list<Cached> cache_;
};

inline void SetCoherence(Coherence c) {
    Runtime &rt = Runtime::get();
    if (c == Coherence::Host) rt.sync_all();
    rt.coherence = c;
}
inline void Sync(const double *h) { Runtime::get().sync(h); }
inline void SyncAll() { Runtime::get().sync_all(); }
inline void Invalidate(const double *h) { Runtime::get().invalidate(h); }
// release the device copy of a host array (e.g. before freeing it)
inline void Forget(const double *h) { Runtime::get().forget(h); }
// 0 selects the corrected ANS y-edge velocity formula, 1 (default) the reference's
inline void SetAnsCompat(bool on) { Runtime::get().ans_compat = on; }

// lattice tables (d2q9.h:160-163, d3q15.h:251-254; D2Q5/D3Q7 with c_s^2 = 1/3)
struct TabD2Q9 {
    static constexpr int kind = LBM_D2Q9, nc = 9, nd = 2;
    static constexpr int cx[nc] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static constexpr int cy[nc] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    static constexpr int cz[nc] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr double ei[nc] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};
struct TabD2Q5 {
    static constexpr int kind = LBM_D2Q5, nc = 5, nd = 2;
    static constexpr int cx[nc] = {0, 1, 0, -1, 0};
    static constexpr int cy[nc] = {0, 0, 1, 0, -1};
    static constexpr int cz[nc] = {0, 0, 0, 0, 0};
    static constexpr double ei[nc] = {1.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
};
struct TabD3Q15 {
    static constexpr int kind = LBM_D3Q15, nc = 15, nd = 3;
    static constexpr int cx[nc] = {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1};
    static constexpr int cy[nc] = {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1};
    static constexpr int cz[nc] = {0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1};
    static constexpr double ei[nc] = {2.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                                      1.0 / 9.0,  1.0 / 9.0,  1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
                                      1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
};
struct TabD3Q27 {  // no reference counterpart: rest, 6 axis, 12 edges (xy, yz, xz), 8 corners
    static constexpr int kind = LBM_D3Q27, nc = 27, nd = 3;
    static constexpr int cx[nc] = {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0,
                                   1, 1, -1, -1, 1, -1, 1, 1, -1, 1, -1, -1};
    static constexpr int cy[nc] = {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, -1, 1, -1, 1, -1,
                                   0, 0, 0, 0, 1, 1, -1, 1, -1, -1, 1, -1};
    static constexpr int cz[nc] = {0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 0, 1, 1, -1, -1,
                                   1, -1, 1, -1, 1, 1, 1, -1, -1, -1, -1, 1};
    static constexpr double ei[nc] = {8.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,
                                      2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
                                      1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
                                      1.0 / 54.0,  1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
                                      1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0};
};
struct TabD3Q7 {
    static constexpr int kind = LBM_D3Q7, nc = 7, nd = 3;
    static constexpr int cx[nc] = {0, 1, 0, 0, -1, 0, 0};
    static constexpr int cy[nc] = {0, 0, 1, 0, 0, -1, 0};
    static constexpr int cz[nc] = {0, 0, 0, 1, 0, 0, -1};
    static constexpr double ei[nc] = {0.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
};

// The lattice classes' common body: grid constants and members of d2q9.h:24-158 / d3q15.h:24-249
template <class T, class Tab>
class Lattice {
    static_assert(std::is_same<T, double>::value, "the device path computes in double");

public:
    static constexpr int nc = Tab::nc, nd = Tab::nd;
    static constexpr const int *cx = Tab::cx;
    static constexpr const int *cy = Tab::cy;
    static constexpr const int *cz = Tab::cz;
    static constexpr const double *ei = Tab::ei;

    Lattice(int _lx, int _ly, int _lz, int _PEid, int _mx, int _my, int _mz)
        : lx(_lx), ly(_ly), lz(_lz), PEid(_PEid), mx(_mx), my(_my), mz(_mz), PEx(0), PEy(0), PEz(0), nx(_lx),
          ny(_ly), nz(_lz), nxyz(_lx * _ly * _lz), offsetx(0), offsety(0), offsetz(0) {
        if (_lx <= 0 || _ly <= 0 || _lz <= 0) throw Error("lattice: sizes must be positive");
        if (_PEid != 0 || _mx != 1 || _my != 1 || _mz != 1)
            throw Error("lattice: one GPU holds the whole domain (PEid = 0, mx = my = mz = 1)");
        check(lbm_lattice_create(Runtime::get().ctx(), Tab::kind, lx, ly, lz, &lat_), "lbm_lattice_create");
        nface_ = lbm_lattice_nface(lat_);
        f0 = new T[nxyz];
        f = new T[(size_t)nxyz * (nc - 1)];
    }
    Lattice(const Lattice &) = delete;
    Lattice &operator=(const Lattice &) = delete;
    ~Lattice() {
        try {
            Runtime::get().lattice_gone(lat_);
        } catch (const Error &) {
            // a destructor cannot report; the lattice is released regardless
        }
        lbm_lattice_destroy(lat_);
        delete[] f0;
        delete[] f;
    }

    int Index(int _i, int _j) const { return _i + nx * _j; }  // d2q9.h:80-84
    int Index(int _i, int _j, int _k) const { return _i + nx * _j + nx * ny * _k; }
    static int IndexF(int _idx, int _c) { return (nc - 1) * _idx + (_c - 1); }  // d2q9.h:88-90

    void Stream() {  // d2q9.h:165-295, d3q15.h:256-617
        Runtime::get().record(lat_, "Stream", [&] { return lbm_stream(lat_); });
    }
    void iStream() {  // d2q9.h:297-427, d3q15.h:619-980
        Runtime::get().record(lat_, "iStream", [&] { return lbm_istream(lat_); });
    }
    template <class Ff>
    void BoundaryCondition(Ff _bctype) {  // d2q9.h:113-119, d3q15.h:182-190
        std::vector<int> type = faces_int(_bctype);
        Runtime::get().record(lat_, "BoundaryCondition", [&] { return lbm_bc(lat_, type.data()); });
    }
    template <class Ff>
    void iBoundaryCondition(Ff _bctype) {  // d2q9.h:120-126, d3q15.h:191-198
        std::vector<int> type = faces_int(_bctype);
        Runtime::get().record(lat_, "iBoundaryCondition", [&] { return lbm_ibc(lat_, type.data()); });
    }
    void SmoothCorner() {  // d2q9.h:127-132, d3q15.h:199-220
        Runtime::get().record(lat_, "SmoothCorner", [&] { return lbm_smooth_corner(lat_); });
    }

    // populations <-> the public host arrays f0, f (reference layout)
    void Download() {
        Runtime::get().flush();
        check(lbm_lattice_download_aos(lat_, f0, f), "lbm_lattice_download_aos");
    }
    void Upload() {
        Runtime::get().flush();
        check(lbm_lattice_upload_aos(lat_, f0, f), "lbm_lattice_upload_aos");
    }

    // ---- face tabulation of the reference's lambdas (global coordinates) ----
    // face order xmin, xmax, ymin, ymax[, zmin, zmax]; 2D x-faces index j, y-faces index i;
    // 3D x-faces j + ly*k, y-faces i + lx*k, z-faces i + lx*j
    template <class Fn>
    void for_faces(Fn fn) const {
        int t = 0;
        if (nd == 2) {
            for (int side = 0; side < 2; ++side)
                for (int j = 0; j < ly; ++j) fn(t++, side ? lx - 1 : 0, j, 0);
            for (int side = 0; side < 2; ++side)
                for (int i = 0; i < lx; ++i) fn(t++, i, side ? ly - 1 : 0, 0);
        } else {
            for (int side = 0; side < 2; ++side)
                for (int k = 0; k < lz; ++k)
                    for (int j = 0; j < ly; ++j) fn(t++, side ? lx - 1 : 0, j, k);
            for (int side = 0; side < 2; ++side)
                for (int k = 0; k < lz; ++k)
                    for (int i = 0; i < lx; ++i) fn(t++, i, side ? ly - 1 : 0, k);
            for (int side = 0; side < 2; ++side)
                for (int j = 0; j < ly; ++j)
                    for (int i = 0; i < lx; ++i) fn(t++, i, j, side ? lz - 1 : 0);
        }
    }
    template <class F>
    auto call(F &fn, int i, int j, int k) const {
        if constexpr (nd == 2) {
            (void)k;
            return fn(i + offsetx, j + offsety);
        } else {
            return fn(i + offsetx, j + offsety, k + offsetz);
        }
    }
    template <class Ff>
    std::vector<int> faces_int(Ff &fn) const {
        std::vector<int> v(nface_, 0);
        for_faces([&](int t, int i, int j, int k) { v[t] = static_cast<int>(call(fn, i, j, k)); });
        return v;
    }
    template <class Fv>
    std::vector<double> faces_val(Fv &fn, const std::vector<int> &mask) const {
        std::vector<double> v(nface_, 0.0);
        for_faces([&](int t, int i, int j, int k) {
            if (mask[t]) v[t] = static_cast<double>(call(fn, i, j, k));
        });
        return v;
    }
    lbm_lattice *handle() const { return lat_; }

    const int lx, ly, lz, PEid, mx, my, mz, PEx, PEy, PEz, nx, ny, nz, nxyz, offsetx, offsety, offsetz;
    T *f0, *f;

private:
    lbm_lattice *lat_ = nullptr;
    int nface_ = 0;
};

}  // namespace gpu

namespace LBM2D {  // the reference's boundary type names (d2q9.h:19-22)
const int BARRIER = LBM_BARRIER;
const int MIRROR = LBM_MIRROR;
}  // namespace LBM2D

template <class T>
class D2Q9 : public gpu::Lattice<T, gpu::TabD2Q9> {
public:
    D2Q9(int _lx, int _ly, int _PEid = 0, int _mx = 1, int _my = 1)
        : gpu::Lattice<T, gpu::TabD2Q9>(_lx, _ly, 1, _PEid, _mx, _my, 1) {}
};
template <class T>
class D2Q5 : public gpu::Lattice<T, gpu::TabD2Q5> {  // no reference counterpart
public:
    D2Q5(int _lx, int _ly, int _PEid = 0, int _mx = 1, int _my = 1)
        : gpu::Lattice<T, gpu::TabD2Q5>(_lx, _ly, 1, _PEid, _mx, _my, 1) {}
};
template <class T>
class D3Q15 : public gpu::Lattice<T, gpu::TabD3Q15> {
public:
    D3Q15(int _lx, int _ly, int _lz, int _PEid = 0, int _mx = 1, int _my = 1, int _mz = 1)
        : gpu::Lattice<T, gpu::TabD3Q15>(_lx, _ly, _lz, _PEid, _mx, _my, _mz) {}
};
template <class T>
class D3Q27 : public gpu::Lattice<T, gpu::TabD3Q27> {  // no reference counterpart: NS with D3Q7 AD
public:
    D3Q27(int _lx, int _ly, int _lz, int _PEid = 0, int _mx = 1, int _my = 1, int _mz = 1)
        : gpu::Lattice<T, gpu::TabD3Q27>(_lx, _ly, _lz, _PEid, _mx, _my, _mz) {}
};
template <class T>
class D3Q7 : public gpu::Lattice<T, gpu::TabD3Q7> {  // no reference counterpart
public:
    D3Q7(int _lx, int _ly, int _lz, int _PEid = 0, int _mx = 1, int _my = 1, int _mz = 1)
        : gpu::Lattice<T, gpu::TabD3Q7>(_lx, _ly, _lz, _PEid, _mx, _my, _mz) {}
};

namespace gpu {
template <class L>
inline size_t N(const L &p) {
    return (size_t)p.nxyz;
}
}  // namespace gpu

// ---- src/utility/residual.h ----------------------------------------------------------------
template <class T>
T Residual(const T *_ux, const T *_uxp, int _nxy) {  // residual.h:9-21
    gpu::Runtime &rt = gpu::Runtime::get();
    double r = 0.0;
    gpu::check(lbm_residual(rt.ctx(), 1, rt.in(_ux, _nxy), nullptr, nullptr, rt.in(_uxp, _nxy), nullptr, nullptr,
                            _nxy, &r),
               "Residual");
    return r;
}
template <class T>
T Residual(const T *_ux, const T *_uy, const T *_uxp, const T *_uyp, int _nxy) {  // residual.h:24-36
    gpu::Runtime &rt = gpu::Runtime::get();
    double r = 0.0;
    gpu::check(lbm_residual(rt.ctx(), 2, rt.in(_ux, _nxy), rt.in(_uy, _nxy), nullptr, rt.in(_uxp, _nxy),
                            rt.in(_uyp, _nxy), nullptr, _nxy, &r),
               "Residual");
    return r;
}
template <class T>
T Residual(const T *_ux, const T *_uy, const T *_uz, const T *_uxp, const T *_uyp, const T *_uzp,
           int _nxyz) {  // residual.h:39-51
    gpu::Runtime &rt = gpu::Runtime::get();
    double r = 0.0;
    gpu::check(lbm_residual(rt.ctx(), 3, rt.in(_ux, _nxyz), rt.in(_uy, _nxyz), rt.in(_uz, _nxyz), rt.in(_uxp, _nxyz),
                            rt.in(_uyp, _nxyz), rt.in(_uzp, _nxyz), _nxyz, &r),
               "Residual");
    return r;
}

// ---- src/utility/normalize.h ---------------------------------------------------------------
template <class T>
void Normalize(T *_v, int _n) {  // normalize.h:9-24
    gpu::Runtime &rt = gpu::Runtime::get();
    double *d = rt.inout(_v, _n);
    gpu::check(lbm_normalize(rt.ctx(), d, _n), "Normalize");
    rt.wrote({_v}, true);
}

// ---- src/equation/navierstokes.h -----------------------------------------------------------
namespace NS {
template <class T, template <class> class P>
void MacroCollide(P<T> &_p, T *_rho, T *_ux, T *_uy, T _viscosity, bool _issave = false) {  // :432-455
    static_assert(P<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n);
    gpu::check(lbm_ns_macro_collide_p(rt.take(), _p.handle(), rho, ux, uy, nullptr, _viscosity, _issave ? 1 : 0),
               "NS::MacroCollide");
    if (_issave) rt.wrote({_rho, _ux, _uy});
}
template <class T, template <class> class P>
void MacroCollide(P<T> &_p, T *_rho, T *_ux, T *_uy, T *_uz, T _viscosity, bool _issave = false) {  // :459-483
    static_assert(P<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n), *uz = rt.out(_uz, n);
    gpu::check(lbm_ns_macro_collide_p(rt.take(), _p.handle(), rho, ux, uy, uz, _viscosity, _issave ? 1 : 0),
               "NS::MacroCollide");
    if (_issave) rt.wrote({_rho, _ux, _uy, _uz});
}
template <class T, template <class> class P>
void MacroBrinkmanCollide(P<T> &_p, T *_rho, T *_ux, T *_uy, T _viscosity, const T *_alpha,
                          bool _issave = false) {  // :487-514
    static_assert(P<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *alpha = rt.in(_alpha, n);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n);
    gpu::check(lbm_ns_macro_brinkman_collide_p(rt.take(), _p.handle(), rho, ux, uy, nullptr, _viscosity, alpha,
                                               _issave ? 1 : 0),
               "NS::MacroBrinkmanCollide");
    if (_issave) rt.wrote({_rho, _ux, _uy});
}
template <class T, template <class> class P>
void MacroBrinkmanCollide(P<T> &_p, T *_rho, T *_ux, T *_uy, T *_uz, T _viscosity, const T *_alpha,
                          bool _issave = false) {  // :518-546
    static_assert(P<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *alpha = rt.in(_alpha, n);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n), *uz = rt.out(_uz, n);
    gpu::check(lbm_ns_macro_brinkman_collide_p(rt.take(), _p.handle(), rho, ux, uy, uz, _viscosity, alpha,
                                               _issave ? 1 : 0),
               "NS::MacroBrinkmanCollide");
    if (_issave) rt.wrote({_rho, _ux, _uy, _uz});
}
template <class T, template <class> class P>
void InitialCondition(P<T> &_p, const T *_rho, const T *_ux, const T *_uy) {  // :550-559
    static_assert(P<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    rt.recheck_host_owned();
    size_t n = gpu::N(_p);
    gpu::check(lbm_ns_initial_condition(_p.handle(), rt.in(_rho, n), rt.in(_ux, n), rt.in(_uy, n), nullptr),
               "NS::InitialCondition");
}
template <class T, template <class> class P>
void InitialCondition(P<T> &_p, const T *_rho, const T *_ux, const T *_uy, const T *_uz) {  // :563-572
    static_assert(P<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    rt.recheck_host_owned();
    size_t n = gpu::N(_p);
    gpu::check(lbm_ns_initial_condition(_p.handle(), rt.in(_rho, n), rt.in(_ux, n), rt.in(_uy, n), rt.in(_uz, n)),
               "NS::InitialCondition");
}
template <class T, template <class> class P, class Fv0, class Fv1, class Ff>
void BoundaryConditionSetU(P<T> &_p, Fv0 _uxbc, Fv1 _uybc, Ff _bctype) {  // :576-581
    static_assert(P<T>::nd == 2, "2D overload");
    std::vector<int> m = _p.faces_int(_bctype);
    std::vector<double> ux = _p.faces_val(_uxbc, m), uy = _p.faces_val(_uybc, m);
    gpu::Runtime::get().record(_p.handle(), "NS::BoundaryConditionSetU", [&] {
        return lbm_ns_set_u(_p.handle(), m.data(), ux.data(), uy.data(), nullptr);
    });
}
template <class T, template <class> class P, class Fv0, class Fv1, class Fv2, class Ff>
void BoundaryConditionSetU(P<T> &_p, Fv0 _uxbc, Fv1 _uybc, Fv2 _uzbc, Ff _bctype) {  // :585-592
    static_assert(P<T>::nd == 3, "3D overload");
    std::vector<int> m = _p.faces_int(_bctype);
    std::vector<double> ux = _p.faces_val(_uxbc, m), uy = _p.faces_val(_uybc, m), uz = _p.faces_val(_uzbc, m);
    gpu::Runtime::get().record(_p.handle(), "NS::BoundaryConditionSetU", [&] {
        return lbm_ns_set_u(_p.handle(), m.data(), ux.data(), uy.data(), uz.data());
    });
}
template <class T, template <class> class P, class Fv0, class Fv1, class Ff>
void BoundaryConditionSetRho(P<T> &_p, Fv0 _rhobc, Fv1 _usbc, Ff _bctype) {  // :596-601
    static_assert(P<T>::nd == 2, "2D overload");
    std::vector<int> m = _p.faces_int(_bctype);
    std::vector<double> rho = _p.faces_val(_rhobc, m), us = _p.faces_val(_usbc, m);
    gpu::Runtime::get().record(_p.handle(), "NS::BoundaryConditionSetRho", [&] {
        return lbm_ns_set_rho(_p.handle(), m.data(), rho.data(), us.data(), nullptr);
    });
}
template <class T, template <class> class P, class Fv0, class Fv1, class Fv2, class Ff>
void BoundaryConditionSetRho(P<T> &_p, Fv0 _rhobc, Fv1 _usbc, Fv2 _utbc, Ff _bctype) {  // :605-612
    static_assert(P<T>::nd == 3, "3D overload");
    std::vector<int> m = _p.faces_int(_bctype);
    std::vector<double> rho = _p.faces_val(_rhobc, m), us = _p.faces_val(_usbc, m), ut = _p.faces_val(_utbc, m);
    gpu::Runtime::get().record(_p.handle(), "NS::BoundaryConditionSetRho", [&] {
        return lbm_ns_set_rho(_p.handle(), m.data(), rho.data(), us.data(), ut.data());
    });
}
}  // namespace NS

// ---- src/equation/advection.h --------------------------------------------------------------
namespace AD {
// :814-867.  _heat (no reference counterpart) adds a volumetric source w_c*heat after the
// Brinkman step; nullptr (the default) is the reference's function.
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideForceConvection(P<T> &_p, T *_rho, T *_ux, T *_uy, const T *_alpha, T _viscosity, Q<T> &_q,
                                         T *_tem, T *_qx, T *_qy, const T *_diffusivity, bool _issave = false,
                                         T *_g = nullptr, const T *_heat = nullptr) {
    static_assert(P<T>::nd == 2 && Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *alpha = rt.in(_alpha, n), *diff = rt.in(_diffusivity, n), *heat = rt.in(_heat, n);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n);
    double *tem = rt.out(_tem, n), *qx = rt.out(_qx, n), *qy = rt.out(_qy, n);
    double *g = _issave ? rt.out(_g, n * Q<T>::nc, Q<T>::nc) : nullptr;
    gpu::check(lbm_ad_macro_brinkman_collide_force_convection_p(rt.take(), _p.handle(), rho, ux, uy, nullptr, alpha,
                                                                _viscosity, _q.handle(), tem, qx, qy, nullptr, diff,
                                                                _issave ? 1 : 0, g, heat),
               "AD::MacroBrinkmanCollideForceConvection");
    if (_issave) rt.wrote({_rho, _ux, _uy, _tem, _qx, _qy, _g});
}
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideForceConvection(P<T> &_p, T *_rho, T *_ux, T *_uy, T *_uz, const T *_alpha, T _viscosity,
                                         Q<T> &_q, T *_tem, T *_qx, T *_qy, T *_qz, const T *_diffusivity,
                                         bool _issave = false, T *_g = nullptr,
                                         const T *_heat = nullptr) {  // :871-926
    static_assert(P<T>::nd == 3 && Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *alpha = rt.in(_alpha, n), *diff = rt.in(_diffusivity, n), *heat = rt.in(_heat, n);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n), *uz = rt.out(_uz, n);
    double *tem = rt.out(_tem, n), *qx = rt.out(_qx, n), *qy = rt.out(_qy, n), *qz = rt.out(_qz, n);
    double *g = _issave ? rt.out(_g, n * Q<T>::nc, Q<T>::nc) : nullptr;
    gpu::check(lbm_ad_macro_brinkman_collide_force_convection_p(rt.take(), _p.handle(), rho, ux, uy, uz, alpha,
                                                                _viscosity, _q.handle(), tem, qx, qy, qz, diff,
                                                                _issave ? 1 : 0, g, heat),
               "AD::MacroBrinkmanCollideForceConvection");
    if (_issave) rt.wrote({_rho, _ux, _uy, _uz, _tem, _qx, _qy, _qz, _g});
}
// :528-569
template <class T, template <class> class P, template <class> class Q>
void MacroCollideForceConvection(P<T> &_p, T *_rho, T *_ux, T *_uy, T _viscosity, Q<T> &_q, T *_tem, T *_qx,
                                 T *_qy, T _diffusivity, bool _issave = false) {
    static_assert(P<T>::nd == 2 && Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n);
    double *tem = rt.out(_tem, n), *qx = rt.out(_qx, n), *qy = rt.out(_qy, n);
    gpu::check(lbm_ad_macro_collide_force_convection_p(rt.take(), _p.handle(), rho, ux, uy, nullptr, _viscosity,
                                                       _q.handle(), tem, qx, qy, nullptr, _diffusivity,
                                                       _issave ? 1 : 0),
               "AD::MacroCollideForceConvection");
    if (_issave) rt.wrote({_rho, _ux, _uy, _tem, _qx, _qy});
}
// :571-614
template <class T, template <class> class P, template <class> class Q>
void MacroCollideForceConvection(P<T> &_p, T *_rho, T *_ux, T *_uy, T *_uz, T _viscosity, Q<T> &_q, T *_tem, T *_qx,
                                 T *_qy, T *_qz, T _diffusivity, bool _issave = false) {
    static_assert(P<T>::nd == 3 && Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n), *uz = rt.out(_uz, n);
    double *tem = rt.out(_tem, n), *qx = rt.out(_qx, n), *qy = rt.out(_qy, n), *qz = rt.out(_qz, n);
    gpu::check(lbm_ad_macro_collide_force_convection_p(rt.take(), _p.handle(), rho, ux, uy, uz, _viscosity,
                                                       _q.handle(), tem, qx, qy, qz, _diffusivity, _issave ? 1 : 0),
               "AD::MacroCollideForceConvection");
    if (_issave) rt.wrote({_rho, _ux, _uy, _uz, _tem, _qx, _qy, _qz});
}
// :616-662
template <class T, template <class> class P, template <class> class Q>
void MacroCollideNaturalConvection(P<T> &_p, T *_rho, T *_ux, T *_uy, T _viscosity, Q<T> &_q, T *_tem, T *_qx,
                                   T *_qy, T _diffusivity, T _gx, T _gy, T _tem0, bool _issave = false) {
    static_assert(P<T>::nd == 2 && Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n);
    double *tem = rt.out(_tem, n), *qx = rt.out(_qx, n), *qy = rt.out(_qy, n);
    gpu::check(lbm_ad_macro_collide_natural_convection_p(rt.take(), _p.handle(), rho, ux, uy, nullptr, _viscosity,
                                                         _q.handle(), tem, qx, qy, nullptr, _diffusivity, _gx, _gy,
                                                         0.0, _tem0, _issave ? 1 : 0),
               "AD::MacroCollideNaturalConvection");
    if (_issave) rt.wrote({_rho, _ux, _uy, _tem, _qx, _qy});
}
// :664-712
template <class T, template <class> class P, template <class> class Q>
void MacroCollideNaturalConvection(P<T> &_p, T *_rho, T *_ux, T *_uy, T *_uz, T _viscosity, Q<T> &_q, T *_tem,
                                   T *_qx, T *_qy, T *_qz, T _diffusivity, T _gx, T _gy, T _gz, T _tem0,
                                   bool _issave = false) {
    static_assert(P<T>::nd == 3 && Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n), *uz = rt.out(_uz, n);
    double *tem = rt.out(_tem, n), *qx = rt.out(_qx, n), *qy = rt.out(_qy, n), *qz = rt.out(_qz, n);
    gpu::check(lbm_ad_macro_collide_natural_convection_p(rt.take(), _p.handle(), rho, ux, uy, uz, _viscosity,
                                                         _q.handle(), tem, qx, qy, qz, _diffusivity, _gx, _gy, _gz,
                                                         _tem0, _issave ? 1 : 0),
               "AD::MacroCollideNaturalConvection");
    if (_issave) rt.wrote({_rho, _ux, _uy, _uz, _tem, _qx, _qy, _qz});
}
// :930-984
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideNaturalConvection(P<T> &_p, T *_rho, T *_ux, T *_uy, const T *_alpha, T _viscosity,
                                           Q<T> &_q, T *_tem, T *_qx, T *_qy, const T *_diffusivity, T _gx, T _gy,
                                           T _tem0, bool _issave = false, T *_g = nullptr) {
    static_assert(P<T>::nd == 2 && Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *alpha = rt.in(_alpha, n), *diff = rt.in(_diffusivity, n);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n);
    double *tem = rt.out(_tem, n), *qx = rt.out(_qx, n), *qy = rt.out(_qy, n);
    double *g = _issave ? rt.out(_g, n * Q<T>::nc, Q<T>::nc) : nullptr;
    gpu::check(lbm_ad_macro_brinkman_collide_natural_convection_p(rt.take(), _p.handle(), rho, ux, uy, nullptr, alpha,
                                                                  _viscosity, _q.handle(), tem, qx, qy, nullptr, diff,
                                                                  _gx, _gy, 0.0, _tem0, _issave ? 1 : 0, g),
               "AD::MacroBrinkmanCollideNaturalConvection");
    if (_issave) rt.wrote({_rho, _ux, _uy, _tem, _qx, _qy, _g});
}
// :988-1044
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideNaturalConvection(P<T> &_p, T *_rho, T *_ux, T *_uy, T *_uz, const T *_alpha, T _viscosity,
                                           Q<T> &_q, T *_tem, T *_qx, T *_qy, T *_qz, const T *_diffusivity, T _gx,
                                           T _gy, T _gz, T _tem0, bool _issave = false, T *_g = nullptr) {
    static_assert(P<T>::nd == 3 && Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *alpha = rt.in(_alpha, n), *diff = rt.in(_diffusivity, n);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n), *uz = rt.out(_uz, n);
    double *tem = rt.out(_tem, n), *qx = rt.out(_qx, n), *qy = rt.out(_qy, n), *qz = rt.out(_qz, n);
    double *g = _issave ? rt.out(_g, n * Q<T>::nc, Q<T>::nc) : nullptr;
    gpu::check(lbm_ad_macro_brinkman_collide_natural_convection_p(rt.take(), _p.handle(), rho, ux, uy, uz, alpha,
                                                                  _viscosity, _q.handle(), tem, qx, qy, qz, diff,
                                                                  _gx, _gy, _gz, _tem0, _issave ? 1 : 0, g),
               "AD::MacroBrinkmanCollideNaturalConvection");
    if (_issave) rt.wrote({_rho, _ux, _uy, _uz, _tem, _qx, _qy, _qz, _g});
}
// :716-760
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideHeatExchange(P<T> &_p, T *_rho, T *_ux, T *_uy, const T *_alpha, T _viscosity, Q<T> &_q,
                                      T *_tem, T *_qx, T *_qy, const T *_beta, T _diffusivity, bool _issave = false) {
    static_assert(P<T>::nd == 2 && Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *alpha = rt.in(_alpha, n), *beta = rt.in(_beta, n);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n);
    double *tem = rt.out(_tem, n), *qx = rt.out(_qx, n), *qy = rt.out(_qy, n);
    gpu::check(lbm_ad_macro_brinkman_collide_heat_exchange_p(rt.take(), _p.handle(), rho, ux, uy, nullptr, alpha,
                                                             _viscosity, _q.handle(), tem, qx, qy, nullptr, beta,
                                                             _diffusivity, _issave ? 1 : 0),
               "AD::MacroBrinkmanCollideHeatExchange");
    if (_issave) rt.wrote({_rho, _ux, _uy, _tem, _qx, _qy});
}
// :764-810
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideHeatExchange(P<T> &_p, T *_rho, T *_ux, T *_uy, T *_uz, const T *_alpha, T _viscosity,
                                      Q<T> &_q, T *_tem, T *_qx, T *_qy, T *_qz, const T *_beta, T _diffusivity,
                                      bool _issave = false) {
    static_assert(P<T>::nd == 3 && Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *alpha = rt.in(_alpha, n), *beta = rt.in(_beta, n);
    double *rho = rt.out(_rho, n), *ux = rt.out(_ux, n), *uy = rt.out(_uy, n), *uz = rt.out(_uz, n);
    double *tem = rt.out(_tem, n), *qx = rt.out(_qx, n), *qy = rt.out(_qy, n), *qz = rt.out(_qz, n);
    gpu::check(lbm_ad_macro_brinkman_collide_heat_exchange_p(rt.take(), _p.handle(), rho, ux, uy, uz, alpha,
                                                             _viscosity, _q.handle(), tem, qx, qy, qz, beta,
                                                             _diffusivity, _issave ? 1 : 0),
               "AD::MacroBrinkmanCollideHeatExchange");
    if (_issave) rt.wrote({_rho, _ux, _uy, _uz, _tem, _qx, _qy, _qz});
}
template <class T, template <class> class Q>
void InitialCondition(Q<T> &_q, const T *_tem, const T *_ux, const T *_uy) {  // :1048-1057
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    rt.recheck_host_owned();
    size_t n = gpu::N(_q);
    gpu::check(lbm_ad_initial_condition(_q.handle(), rt.in(_tem, n), rt.in(_ux, n), rt.in(_uy, n), nullptr),
               "AD::InitialCondition");
}
template <class T, template <class> class Q>
void InitialCondition(Q<T> &_q, const T *_tem, const T *_ux, const T *_uy, const T *_uz) {  // :1061-1070
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    rt.recheck_host_owned();
    size_t n = gpu::N(_q);
    gpu::check(lbm_ad_initial_condition(_q.handle(), rt.in(_tem, n), rt.in(_ux, n), rt.in(_uy, n), rt.in(_uz, n)),
               "AD::InitialCondition");
}
template <class T, template <class> class Q, class Fv, class Ff>
void BoundaryConditionSetT(Q<T> &_q, Fv _tembc, const T *_ux, const T *_uy, Ff _bctype) {  // :1074-1079
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    std::vector<int> m = _q.faces_int(_bctype);
    std::vector<double> tem = _q.faces_val(_tembc, m);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n);
    rt.record(_q.handle(), "AD::BoundaryConditionSetT",
              [&] { return lbm_ad_set_t(_q.handle(), m.data(), tem.data(), ux, uy, nullptr); });
}
template <class T, template <class> class Q, class Fv, class Ff>
void BoundaryConditionSetT(Q<T> &_q, Fv _tembc, const T *_ux, const T *_uy, const T *_uz,
                           Ff _bctype) {  // :1083-1090
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    std::vector<int> m = _q.faces_int(_bctype);
    std::vector<double> tem = _q.faces_val(_tembc, m);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    rt.record(_q.handle(), "AD::BoundaryConditionSetT",
              [&] { return lbm_ad_set_t(_q.handle(), m.data(), tem.data(), ux, uy, uz); });
}
template <class T, template <class> class Q, class Fv, class Ff>
void BoundaryConditionSetQ(Q<T> &_q, Fv _qnbc, const T *_ux, const T *_uy, T _diffusivity,
                           Ff _bctype) {  // :1094-1099
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    std::vector<int> m = _q.faces_int(_bctype);
    std::vector<double> qn = _q.faces_val(_qnbc, m);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n);
    rt.record(_q.handle(), "AD::BoundaryConditionSetQ", [&] {
        return lbm_ad_set_q(_q.handle(), m.data(), qn.data(), ux, uy, nullptr, nullptr, _diffusivity);
    });
}
template <class T, template <class> class Q, class Fv, class Ff>
void BoundaryConditionSetQ(Q<T> &_q, Fv _qnbc, const T *_ux, const T *_uy, const T *_diffusivity,
                           Ff _bctype) {  // :1114-1119
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    std::vector<int> m = _q.faces_int(_bctype);
    std::vector<double> qn = _q.faces_val(_qnbc, m);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *diff = rt.in(_diffusivity, n);
    rt.record(_q.handle(), "AD::BoundaryConditionSetQ", [&] {
        return lbm_ad_set_q(_q.handle(), m.data(), qn.data(), ux, uy, nullptr, diff, 0.0);
    });
}
template <class T, template <class> class Q, class Fv, class Ff>
void BoundaryConditionSetQ(Q<T> &_q, Fv _qnbc, const T *_ux, const T *_uy, const T *_uz, T _diffusivity,
                           Ff _bctype) {  // :1103-1110
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    std::vector<int> m = _q.faces_int(_bctype);
    std::vector<double> qn = _q.faces_val(_qnbc, m);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    rt.record(_q.handle(), "AD::BoundaryConditionSetQ", [&] {
        return lbm_ad_set_q(_q.handle(), m.data(), qn.data(), ux, uy, uz, nullptr, _diffusivity);
    });
}
template <class T, template <class> class Q, class Fv, class Ff>
void BoundaryConditionSetQ(Q<T> &_q, Fv _qnbc, const T *_ux, const T *_uy, const T *_uz, const T *_diffusivity,
                           Ff _bctype) {  // :1123-1130
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    std::vector<int> m = _q.faces_int(_bctype);
    std::vector<double> qn = _q.faces_val(_qnbc, m);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n), *diff = rt.in(_diffusivity, n);
    rt.record(_q.handle(), "AD::BoundaryConditionSetQ", [&] {
        return lbm_ad_set_q(_q.handle(), m.data(), qn.data(), ux, uy, uz, diff, 0.0);
    });
}
}  // namespace AD

// ---- src/equation/adjointnavierstokes.h ----------------------------------------------------
namespace ANS {
template <class T, template <class> class P>
void MacroBrinkmanCollide(P<T> &_p, const T *_rho, const T *_ux, const T *_uy, T *_ip, T *_iux, T *_iuy, T *_imx,
                          T *_imy, T _viscosity, const T *_alpha, bool _issave = false) {  // :398-431
    static_assert(P<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *alpha = rt.in(_alpha, n);
    double *ip = rt.out(_ip, n), *iux = rt.out(_iux, n), *iuy = rt.out(_iuy, n);
    double *imx = rt.out(_imx, n), *imy = rt.out(_imy, n);
    gpu::check(lbm_ans_macro_brinkman_collide_p(rt.take(), _p.handle(), rho, ux, uy, nullptr, ip, iux, iuy, nullptr,
                                                imx, imy, nullptr, _viscosity, alpha, _issave ? 1 : 0),
               "ANS::MacroBrinkmanCollide");
    if (_issave) rt.wrote({_ip, _iux, _iuy, _imx, _imy});
}
template <class T, template <class> class P>
void MacroBrinkmanCollide(P<T> &_p, const T *_rho, const T *_ux, const T *_uy, const T *_uz, T *_ip, T *_iux,
                          T *_iuy, T *_iuz, T *_imx, T *_imy, T *_imz, T _viscosity, const T *_alpha,
                          bool _issave = false) {  // :435-470
    static_assert(P<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    const double *alpha = rt.in(_alpha, n);
    double *ip = rt.out(_ip, n), *iux = rt.out(_iux, n), *iuy = rt.out(_iuy, n), *iuz = rt.out(_iuz, n);
    double *imx = rt.out(_imx, n), *imy = rt.out(_imy, n), *imz = rt.out(_imz, n);
    gpu::check(lbm_ans_macro_brinkman_collide_p(rt.take(), _p.handle(), rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy,
                                                imz, _viscosity, alpha, _issave ? 1 : 0),
               "ANS::MacroBrinkmanCollide");
    if (_issave) rt.wrote({_ip, _iux, _iuy, _iuz, _imx, _imy, _imz});
}
template <class T, template <class> class P>
void InitialCondition(P<T> &_p, const T *_ux, const T *_uy, const T *_ip, const T *_iux,
                      const T *_iuy) {  // :474-484
    static_assert(P<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    rt.recheck_host_owned();
    size_t n = gpu::N(_p);
    gpu::check(lbm_ans_initial_condition(_p.handle(), rt.in(_ux, n), rt.in(_uy, n), nullptr, rt.in(_ip, n),
                                         rt.in(_iux, n), rt.in(_iuy, n), nullptr),
               "ANS::InitialCondition");
}
template <class T, template <class> class P>
void InitialCondition(P<T> &_p, const T *_ux, const T *_uy, const T *_uz, const T *_ip, const T *_iux, const T *_iuy,
                      const T *_iuz) {  // :488-498
    static_assert(P<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    rt.recheck_host_owned();
    size_t n = gpu::N(_p);
    gpu::check(lbm_ans_initial_condition(_p.handle(), rt.in(_ux, n), rt.in(_uy, n), rt.in(_uz, n), rt.in(_ip, n),
                                         rt.in(_iux, n), rt.in(_iuy, n), rt.in(_iuz, n)),
               "ANS::InitialCondition");
}
template <class T, template <class> class P, class Fv0, class Fv1, class Ff>
void iBoundaryConditionSetU(P<T> &_p, Fv0 _uxbc, Fv1 _uybc, Ff _bctype, T _eps = T()) {  // :502-507
    static_assert(P<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    std::vector<int> m = _p.faces_int(_bctype);
    std::vector<double> ux = _p.faces_val(_uxbc, m), uy = _p.faces_val(_uybc, m);
    int compat = rt.ans_compat ? 1 : 0;
    rt.record(_p.handle(), "ANS::iBoundaryConditionSetU", [&] {
        return lbm_ans_iset_u(_p.handle(), m.data(), ux.data(), uy.data(), nullptr, _eps, compat);
    });
}
template <class T, template <class> class P, class Fv0, class Fv1, class Fv2, class Ff>
void iBoundaryConditionSetU(P<T> &_p, Fv0 _uxbc, Fv1 _uybc, Fv2 _uzbc, Ff _bctype, T _eps = T()) {  // :511-518
    static_assert(P<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    std::vector<int> m = _p.faces_int(_bctype);
    std::vector<double> ux = _p.faces_val(_uxbc, m), uy = _p.faces_val(_uybc, m), uz = _p.faces_val(_uzbc, m);
    int compat = rt.ans_compat ? 1 : 0;
    rt.record(_p.handle(), "ANS::iBoundaryConditionSetU", [&] {
        return lbm_ans_iset_u(_p.handle(), m.data(), ux.data(), uy.data(), uz.data(), _eps, compat);
    });
}
template <class T, template <class> class P, class Ff>
void iBoundaryConditionSetRho2D(P<T> &_p, Ff _bctype) {  // :522-527
    static_assert(P<T>::nd == 2, "2D overload");
    std::vector<int> m = _p.faces_int(_bctype);
    gpu::Runtime::get().record(_p.handle(), "ANS::iBoundaryConditionSetRho2D",
                               [&] { return lbm_ans_iset_rho(_p.handle(), m.data()); });
}
template <class T, template <class> class P, class Ff>
void iBoundaryConditionSetRho3D(P<T> &_p, Ff _bctype) {  // :531-538
    static_assert(P<T>::nd == 3, "3D overload");
    std::vector<int> m = _p.faces_int(_bctype);
    gpu::Runtime::get().record(_p.handle(), "ANS::iBoundaryConditionSetRho3D",
                               [&] { return lbm_ans_iset_rho(_p.handle(), m.data()); });
}
template <class T, template <class> class P>
void SensitivityBrinkman(P<T> &_p, T *_dfds, const T *_ux, const T *_uy, const T *_imx, const T *_imy,
                         const T *_dads) {  // :542-546
    static_assert(P<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    size_t n = gpu::N(_p);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *imx = rt.in(_imx, n), *imy = rt.in(_imy, n);
    const double *dads = rt.in(_dads, n);
    double *dfds = rt.inout(_dfds, n);
    gpu::check(lbm_sensitivity_brinkman(_p.handle(), dfds, ux, uy, nullptr, imx, imy, nullptr, dads),
               "ANS::SensitivityBrinkman");
    rt.wrote({_dfds}, true);
}
template <class T, template <class> class P>
void SensitivityBrinkman(P<T> &_p, T *_dfds, const T *_ux, const T *_uy, const T *_uz, const T *_imx, const T *_imy,
                         const T *_imz, const T *_dads) {  // :550-554
    static_assert(P<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    size_t n = gpu::N(_p);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    const double *imx = rt.in(_imx, n), *imy = rt.in(_imy, n), *imz = rt.in(_imz, n), *dads = rt.in(_dads, n);
    double *dfds = rt.inout(_dfds, n);
    gpu::check(lbm_sensitivity_brinkman(_p.handle(), dfds, ux, uy, uz, imx, imy, imz, dads),
               "ANS::SensitivityBrinkman");
    rt.wrote({_dfds}, true);
}
}  // namespace ANS

// ---- src/equation/adjointadvection.h -------------------------------------------------------
namespace AAD {
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideForceConvection(P<T> &_p, const T *_rho, const T *_ux, const T *_uy, T *_ip, T *_iux,
                                         T *_iuy, T *_imx, T *_imy, const T *_alpha, T _viscosity, Q<T> &_q,
                                         const T *_tem, T *_item, T *_iqx, T *_iqy, const T *_diffusivity,
                                         bool _issave = false, T *_ig = nullptr) {  // :997-1050
    static_assert(P<T>::nd == 2 && Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *alpha = rt.in(_alpha, n);
    const double *tem = rt.in(_tem, n), *diff = rt.in(_diffusivity, n);
    double *ip = rt.out(_ip, n), *iux = rt.out(_iux, n), *iuy = rt.out(_iuy, n);
    double *imx = rt.out(_imx, n), *imy = rt.out(_imy, n);
    double *item = rt.out(_item, n), *iqx = rt.out(_iqx, n), *iqy = rt.out(_iqy, n);
    double *ig = _issave ? rt.out(_ig, n * Q<T>::nc, Q<T>::nc) : nullptr;
    gpu::check(lbm_aad_macro_brinkman_collide_force_convection_p(
                   rt.take(), _p.handle(), rho, ux, uy, nullptr, ip, iux, iuy, nullptr, imx, imy, nullptr, alpha,
                   _viscosity, _q.handle(), tem, item, iqx, iqy, nullptr, diff, _issave ? 1 : 0, ig),
               "AAD::MacroBrinkmanCollideForceConvection");
    if (_issave) rt.wrote({_ip, _iux, _iuy, _imx, _imy, _item, _iqx, _iqy, _ig});
}
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideForceConvection(P<T> &_p, const T *_rho, const T *_ux, const T *_uy, const T *_uz, T *_ip,
                                         T *_iux, T *_iuy, T *_iuz, T *_imx, T *_imy, T *_imz, const T *_alpha,
                                         T _viscosity, Q<T> &_q, const T *_tem, T *_item, T *_iqx, T *_iqy, T *_iqz,
                                         const T *_diffusivity, bool _issave = false,
                                         T *_ig = nullptr) {  // :1054-1109
    static_assert(P<T>::nd == 3 && Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    const double *alpha = rt.in(_alpha, n), *tem = rt.in(_tem, n), *diff = rt.in(_diffusivity, n);
    (void)_iuz;  // the reference's 3D save writes no iuz (:1074-1084), so the host array is left alone
    double *ip = rt.out(_ip, n), *iux = rt.out(_iux, n), *iuy = rt.out(_iuy, n);
    double *imx = rt.out(_imx, n), *imy = rt.out(_imy, n), *imz = rt.out(_imz, n);
    double *item = rt.out(_item, n), *iqx = rt.out(_iqx, n), *iqy = rt.out(_iqy, n), *iqz = rt.out(_iqz, n);
    double *ig = _issave ? rt.out(_ig, n * Q<T>::nc, Q<T>::nc) : nullptr;
    gpu::check(lbm_aad_macro_brinkman_collide_force_convection_p(rt.take(), _p.handle(), rho, ux, uy, uz, ip, iux,
                                                                 iuy, nullptr, imx, imy, imz, alpha, _viscosity,
                                                                 _q.handle(), tem, item, iqx, iqy, iqz, diff,
                                                                 _issave ? 1 : 0, ig),
               "AAD::MacroBrinkmanCollideForceConvection");
    if (_issave) rt.wrote({_ip, _iux, _iuy, _imx, _imy, _imz, _item, _iqx, _iqy, _iqz, _ig});
}
// :1113-1168
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideNaturalConvection(P<T> &_p, const T *_rho, const T *_ux, const T *_uy, T *_ip, T *_iux,
                                           T *_iuy, T *_imx, T *_imy, const T *_alpha, T _viscosity, Q<T> &_q,
                                           const T *_tem, T *_item, T *_iqx, T *_iqy, const T *_diffusivity, T _gx,
                                           T _gy, bool _issave = false, T *_ig = nullptr) {
    static_assert(P<T>::nd == 2 && Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *alpha = rt.in(_alpha, n);
    const double *tem = rt.in(_tem, n), *diff = rt.in(_diffusivity, n);
    double *ip = rt.out(_ip, n), *iux = rt.out(_iux, n), *iuy = rt.out(_iuy, n);
    double *imx = rt.out(_imx, n), *imy = rt.out(_imy, n);
    double *item = rt.out(_item, n), *iqx = rt.out(_iqx, n), *iqy = rt.out(_iqy, n);
    double *ig = _issave ? rt.out(_ig, n * Q<T>::nc, Q<T>::nc) : nullptr;
    gpu::check(lbm_aad_macro_brinkman_collide_natural_convection_p(
                   rt.take(), _p.handle(), rho, ux, uy, nullptr, ip, iux, iuy, nullptr, imx, imy, nullptr, alpha,
                   _viscosity, _q.handle(), tem, item, iqx, iqy, nullptr, diff, _gx, _gy, 0.0, _issave ? 1 : 0, ig),
               "AAD::MacroBrinkmanCollideNaturalConvection");
    if (_issave) rt.wrote({_ip, _iux, _iuy, _imx, _imy, _item, _iqx, _iqy, _ig});
}
// :1172-1230
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideNaturalConvection(P<T> &_p, const T *_rho, const T *_ux, const T *_uy, const T *_uz,
                                           T *_ip, T *_iux, T *_iuy, T *_iuz, T *_imx, T *_imy, T *_imz,
                                           const T *_alpha, T _viscosity, Q<T> &_q, const T *_tem, T *_item, T *_iqx,
                                           T *_iqy, T *_iqz, const T *_diffusivity, T _gx, T _gy, T _gz,
                                           bool _issave = false, T *_ig = nullptr) {
    static_assert(P<T>::nd == 3 && Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    const double *alpha = rt.in(_alpha, n), *tem = rt.in(_tem, n), *diff = rt.in(_diffusivity, n);
    double *ip = rt.out(_ip, n), *iux = rt.out(_iux, n), *iuy = rt.out(_iuy, n), *iuz = rt.out(_iuz, n);
    double *imx = rt.out(_imx, n), *imy = rt.out(_imy, n), *imz = rt.out(_imz, n);
    double *item = rt.out(_item, n), *iqx = rt.out(_iqx, n), *iqy = rt.out(_iqy, n), *iqz = rt.out(_iqz, n);
    double *ig = _issave ? rt.out(_ig, n * Q<T>::nc, Q<T>::nc) : nullptr;
    gpu::check(lbm_aad_macro_brinkman_collide_natural_convection_p(
                   rt.take(), _p.handle(), rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha, _viscosity,
                   _q.handle(), tem, item, iqx, iqy, iqz, diff, _gx, _gy, _gz, _issave ? 1 : 0, ig),
               "AAD::MacroBrinkmanCollideNaturalConvection");
    if (_issave) rt.wrote({_ip, _iux, _iuy, _iuz, _imx, _imy, _imz, _item, _iqx, _iqy, _iqz, _ig});
}
// :1232-1291
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideNaturalConvectionMassFlow(P<T> &_p, const T *_rho, const T *_ux, const T *_uy, T *_ip,
                                                   T *_iux, T *_iuy, T *_imx, T *_imy, const T *_alpha, T _viscosity,
                                                   Q<T> &_q, const T *_tem, T *_item, T *_iqx, T *_iqy,
                                                   const T *_diffusivity, T _gx, T _gy, const T *_directionx,
                                                   const T *_directiony, bool _issave = false, T *_ig = nullptr) {
    static_assert(P<T>::nd == 2 && Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *alpha = rt.in(_alpha, n);
    const double *tem = rt.in(_tem, n), *diff = rt.in(_diffusivity, n);
    const double *dx = rt.in(_directionx, n), *dy = rt.in(_directiony, n);
    double *ip = rt.out(_ip, n), *iux = rt.out(_iux, n), *iuy = rt.out(_iuy, n);
    double *imx = rt.out(_imx, n), *imy = rt.out(_imy, n);
    double *item = rt.out(_item, n), *iqx = rt.out(_iqx, n), *iqy = rt.out(_iqy, n);
    double *ig = _issave ? rt.out(_ig, n * Q<T>::nc, Q<T>::nc) : nullptr;
    gpu::check(lbm_aad_macro_brinkman_collide_natural_convection_mass_flow_p(
                   rt.take(), _p.handle(), rho, ux, uy, nullptr, ip, iux, iuy, nullptr, imx, imy, nullptr, alpha,
                   _viscosity, _q.handle(), tem, item, iqx, iqy, nullptr, diff, _gx, _gy, 0.0, dx, dy, nullptr,
                   _issave ? 1 : 0, ig),
               "AAD::MacroBrinkmanCollideNaturalConvectionMassFlow");
    if (_issave) rt.wrote({_ip, _iux, _iuy, _imx, _imy, _item, _iqx, _iqy, _ig});
}
// :1293-1358
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideNaturalConvectionMassFlow(P<T> &_p, const T *_rho, const T *_ux, const T *_uy,
                                                   const T *_uz, T *_ip, T *_iux, T *_iuy, T *_iuz, T *_imx, T *_imy,
                                                   T *_imz, const T *_alpha, T _viscosity, Q<T> &_q, const T *_tem,
                                                   T *_item, T *_iqx, T *_iqy, T *_iqz, const T *_diffusivity, T _gx,
                                                   T _gy, T _gz, const T *_directionx, const T *_directiony,
                                                   const T *_directionz, bool _issave = false, T *_ig = nullptr) {
    static_assert(P<T>::nd == 3 && Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    const double *alpha = rt.in(_alpha, n), *tem = rt.in(_tem, n), *diff = rt.in(_diffusivity, n);
    const double *dx = rt.in(_directionx, n), *dy = rt.in(_directiony, n), *dz = rt.in(_directionz, n);
    double *ip = rt.out(_ip, n), *iux = rt.out(_iux, n), *iuy = rt.out(_iuy, n), *iuz = rt.out(_iuz, n);
    double *imx = rt.out(_imx, n), *imy = rt.out(_imy, n), *imz = rt.out(_imz, n);
    double *item = rt.out(_item, n), *iqx = rt.out(_iqx, n), *iqy = rt.out(_iqy, n), *iqz = rt.out(_iqz, n);
    double *ig = _issave ? rt.out(_ig, n * Q<T>::nc, Q<T>::nc) : nullptr;
    gpu::check(lbm_aad_macro_brinkman_collide_natural_convection_mass_flow_p(
                   rt.take(), _p.handle(), rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha, _viscosity,
                   _q.handle(), tem, item, iqx, iqy, iqz, diff, _gx, _gy, _gz, dx, dy, dz, _issave ? 1 : 0, ig),
               "AAD::MacroBrinkmanCollideNaturalConvectionMassFlow");
    if (_issave) rt.wrote({_ip, _iux, _iuy, _iuz, _imx, _imy, _imz, _item, _iqx, _iqy, _iqz, _ig});
}
template <class T, template <class> class Q>
void InitialCondition(Q<T> &_q, const T *_ux, const T *_uy, const T *_item, const T *_iqx,
                      const T *_iqy) {  // :1359-1368
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    rt.recheck_host_owned();
    size_t n = gpu::N(_q);
    gpu::check(lbm_aad_initial_condition(_q.handle(), rt.in(_ux, n), rt.in(_uy, n), nullptr, rt.in(_item, n),
                                         rt.in(_iqx, n), rt.in(_iqy, n), nullptr),
               "AAD::InitialCondition");
}
template <class T, template <class> class Q>
void InitialCondition(Q<T> &_q, const T *_ux, const T *_uy, const T *_uz, const T *_item, const T *_iqx,
                      const T *_iqy, const T *_iqz) {  // :1372-1381
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    rt.recheck_host_owned();
    size_t n = gpu::N(_q);
    gpu::check(lbm_aad_initial_condition(_q.handle(), rt.in(_ux, n), rt.in(_uy, n), rt.in(_uz, n), rt.in(_item, n),
                                         rt.in(_iqx, n), rt.in(_iqy, n), rt.in(_iqz, n)),
               "AAD::InitialCondition");
}
template <class T, template <class> class Q, class Ff>
void iBoundaryConditionSetT(Q<T> &_q, const T *_ux, const T *_uy, Ff _bctype) {  // :1385-1390
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    std::vector<int> m = _q.faces_int(_bctype);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n);
    rt.record(_q.handle(), "AAD::iBoundaryConditionSetT",
              [&] { return lbm_aad_iset_t(_q.handle(), m.data(), ux, uy, nullptr); });
}
template <class T, template <class> class Q, class Ff>
void iBoundaryConditionSetT(Q<T> &_q, const T *_ux, const T *_uy, const T *_uz, Ff _bctype) {  // :1394-1401
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    std::vector<int> m = _q.faces_int(_bctype);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    rt.record(_q.handle(), "AAD::iBoundaryConditionSetT",
              [&] { return lbm_aad_iset_t(_q.handle(), m.data(), ux, uy, uz); });
}
template <class T, template <class> class Q, class Ff>
void iBoundaryConditionSetQ(Q<T> &_q, const T *_ux, const T *_uy, Ff _bctype, T _eps = T()) {  // :1405-1410
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    std::vector<int> m = _q.faces_int(_bctype);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n);
    rt.record(_q.handle(), "AAD::iBoundaryConditionSetQ",
              [&] { return lbm_aad_iset_q(_q.handle(), m.data(), ux, uy, nullptr, _eps); });
}
template <class T, template <class> class Q, class Ff>
void iBoundaryConditionSetQ(Q<T> &_q, const T *_ux, const T *_uy, const T *_uz, Ff _bctype,
                            T _eps = T()) {  // :1414-1421
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    std::vector<int> m = _q.faces_int(_bctype);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    rt.record(_q.handle(), "AAD::iBoundaryConditionSetQ",
              [&] { return lbm_aad_iset_q(_q.handle(), m.data(), ux, uy, uz, _eps); });
}
template <class T, template <class> class P, template <class> class Q, class Ff>
void iBoundaryConditionSetRho(P<T> &_p, Q<T> &_q, const T *_rho, const T *_ux, const T *_uy, const T *_tem,
                              Ff _bctype, T _eps = T()) {  // :1425-1430
    static_assert(P<T>::nd == 2 && Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    std::vector<int> m = _p.faces_int(_bctype);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *tem = rt.in(_tem, n);
    rt.record(
        _p.handle(), "AAD::iBoundaryConditionSetRho",
        [&] { return lbm_aad_iset_rho(_p.handle(), _q.handle(), m.data(), rho, ux, uy, nullptr, tem, _eps); },
        _q.handle());
}
// :1445-1450
template <class T, template <class> class Q>
void SensitivityHeatExchange(Q<T> &_q, T *_dfds, const T *_ux, const T *_uy, const T *_imx, const T *_imy,
                             const T *_dads, const T *_tem, const T *_item, const T *_dbds) {
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    double *dfds = rt.inout(_dfds, n);
    gpu::check(lbm_sensitivity_heat_exchange(_q.handle(), dfds, rt.in(_ux, n), rt.in(_uy, n), nullptr, rt.in(_imx, n),
                                             rt.in(_imy, n), nullptr, rt.in(_dads, n), rt.in(_tem, n),
                                             rt.in(_item, n), rt.in(_dbds, n)),
               "AAD::SensitivityHeatExchange");
    rt.wrote({_dfds}, true);
}
// :1452-1457
template <class T, template <class> class Q>
void SensitivityHeatExchange(Q<T> &_q, T *_dfds, const T *_ux, const T *_uy, const T *_uz, const T *_imx,
                             const T *_imy, const T *_imz, const T *_dads, const T *_tem, const T *_item,
                             const T *_dbds) {
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    double *dfds = rt.inout(_dfds, n);
    gpu::check(lbm_sensitivity_heat_exchange(_q.handle(), dfds, rt.in(_ux, n), rt.in(_uy, n), rt.in(_uz, n),
                                             rt.in(_imx, n), rt.in(_imy, n), rt.in(_imz, n), rt.in(_dads, n),
                                             rt.in(_tem, n), rt.in(_item, n), rt.in(_dbds, n)),
               "AAD::SensitivityHeatExchange");
    rt.wrote({_dfds}, true);
}
// No reference counterpart: the design derivative of the volumetric source that the _heat argument
// of AD::MacroBrinkmanCollideForceConvection adds (lbm_sensitivity_heat_source):
// dfds += -dhds (item + 3 omega_g u.iq), omega_g = 1/(3 D + 1/2).
template <class T, template <class> class Q>
void SensitivityHeatSource(Q<T> &_q, T *_dfds, const T *_ux, const T *_uy, const T *_item, const T *_iqx,
                           const T *_iqy, const T *_diffusivity, const T *_dhds) {
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    double *dfds = rt.inout(_dfds, n);
    gpu::check(lbm_sensitivity_heat_source(_q.handle(), dfds, rt.in(_ux, n), rt.in(_uy, n), nullptr, rt.in(_item, n),
                                           rt.in(_iqx, n), rt.in(_iqy, n), nullptr, rt.in(_diffusivity, n),
                                           rt.in(_dhds, n)),
               "AAD::SensitivityHeatSource");
    rt.wrote({_dfds}, true);
}
template <class T, template <class> class Q>
void SensitivityHeatSource(Q<T> &_q, T *_dfds, const T *_ux, const T *_uy, const T *_uz, const T *_item,
                           const T *_iqx, const T *_iqy, const T *_iqz, const T *_diffusivity, const T *_dhds) {
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_q);
    double *dfds = rt.inout(_dfds, n);
    gpu::check(lbm_sensitivity_heat_source(_q.handle(), dfds, rt.in(_ux, n), rt.in(_uy, n), rt.in(_uz, n),
                                           rt.in(_item, n), rt.in(_iqx, n), rt.in(_iqy, n), rt.in(_iqz, n),
                                           rt.in(_diffusivity, n), rt.in(_dhds, n)),
               "AAD::SensitivityHeatSource");
    rt.wrote({_dfds}, true);
}
// :894-940
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideHeatExchange(P<T> &_p, const T *_rho, const T *_ux, const T *_uy, T *_ip, T *_iux, T *_iuy,
                                      T *_imx, T *_imy, const T *_alpha, T _viscosity, Q<T> &_q, const T *_tem,
                                      T *_item, T *_iqx, T *_iqy, const T *_beta, T _diffusivity,
                                      bool _issave = false) {
    static_assert(P<T>::nd == 2 && Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *alpha = rt.in(_alpha, n);
    const double *tem = rt.in(_tem, n), *beta = rt.in(_beta, n);
    double *ip = rt.out(_ip, n), *iux = rt.out(_iux, n), *iuy = rt.out(_iuy, n);
    double *imx = rt.out(_imx, n), *imy = rt.out(_imy, n);
    double *item = rt.out(_item, n), *iqx = rt.out(_iqx, n), *iqy = rt.out(_iqy, n);
    gpu::check(lbm_aad_macro_brinkman_collide_heat_exchange_p(
                   rt.take(), _p.handle(), rho, ux, uy, nullptr, ip, iux, iuy, nullptr, imx, imy, nullptr, alpha,
                   _viscosity, _q.handle(), tem, item, iqx, iqy, nullptr, beta, _diffusivity, _issave ? 1 : 0),
               "AAD::MacroBrinkmanCollideHeatExchange");
    if (_issave) rt.wrote({_ip, _iux, _iuy, _imx, _imy, _item, _iqx, _iqy});
}
// :944-993
template <class T, template <class> class P, template <class> class Q>
void MacroBrinkmanCollideHeatExchange(P<T> &_p, const T *_rho, const T *_ux, const T *_uy, const T *_uz, T *_ip,
                                      T *_iux, T *_iuy, T *_iuz, T *_imx, T *_imy, T *_imz, const T *_alpha,
                                      T _viscosity, Q<T> &_q, const T *_tem, T *_item, T *_iqx, T *_iqy, T *_iqz,
                                      const T *_beta, T _diffusivity, bool _issave = false) {
    static_assert(P<T>::nd == 3 && Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    size_t n = gpu::N(_p);
    const double *rho = rt.in(_rho, n), *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    const double *alpha = rt.in(_alpha, n), *tem = rt.in(_tem, n), *beta = rt.in(_beta, n);
    double *ip = rt.out(_ip, n), *iux = rt.out(_iux, n), *iuy = rt.out(_iuy, n), *iuz = rt.out(_iuz, n);
    double *imx = rt.out(_imx, n), *imy = rt.out(_imy, n), *imz = rt.out(_imz, n);
    double *item = rt.out(_item, n), *iqx = rt.out(_iqx, n), *iqy = rt.out(_iqy, n), *iqz = rt.out(_iqz, n);
    gpu::check(lbm_aad_macro_brinkman_collide_heat_exchange_p(rt.take(), _p.handle(), rho, ux, uy, uz, ip, iux, iuy,
                                                              iuz, imx, imy, imz, alpha, _viscosity, _q.handle(), tem,
                                                              item, iqx, iqy, iqz, beta, _diffusivity,
                                                              _issave ? 1 : 0),
               "AAD::MacroBrinkmanCollideHeatExchange");
    if (_issave) rt.wrote({_ip, _iux, _iuy, _iuz, _imx, _imy, _imz, _item, _iqx, _iqy, _iqz});
}
template <class T, template <class> class Q>
void SensitivityBrinkmanDiffusivity(Q<T> &_q, T *_dfds, const T *_ux, const T *_uy, const T *_imx, const T *_imy,
                                    const T *_dads, const T *_tem, const T *_item, const T *_iqx, const T *_iqy,
                                    const T *_g, const T *_ig, const T *_diffusivity,
                                    const T *_dkds) {  // :1461-1476
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    size_t n = gpu::N(_q), ng = n * Q<T>::nc;
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *imx = rt.in(_imx, n), *imy = rt.in(_imy, n);
    const double *dads = rt.in(_dads, n), *tem = rt.in(_tem, n), *item = rt.in(_item, n);
    const double *iqx = rt.in(_iqx, n), *iqy = rt.in(_iqy, n);
    const double *g = rt.in(_g, ng, Q<T>::nc), *ig = rt.in(_ig, ng, Q<T>::nc);
    const double *diff = rt.in(_diffusivity, n), *dkds = rt.in(_dkds, n);
    double *dfds = rt.inout(_dfds, n);
    gpu::check(lbm_sensitivity_brinkman_diffusivity(_q.handle(), dfds, ux, uy, nullptr, imx, imy, nullptr, dads, tem,
                                                    item, iqx, iqy, nullptr, g, ig, diff, dkds),
               "AAD::SensitivityBrinkmanDiffusivity");
    rt.wrote({_dfds}, true);
}
template <class T, template <class> class Q>
void SensitivityBrinkmanDiffusivity(Q<T> &_q, T *_dfds, const T *_ux, const T *_uy, const T *_uz, const T *_imx,
                                    const T *_imy, const T *_imz, const T *_dads, const T *_tem, const T *_item,
                                    const T *_iqx, const T *_iqy, const T *_iqz, const T *_g, const T *_ig,
                                    const T *_diffusivity, const T *_dkds) {  // :1478-1493
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    size_t n = gpu::N(_q), ng = n * Q<T>::nc;
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    const double *imx = rt.in(_imx, n), *imy = rt.in(_imy, n), *imz = rt.in(_imz, n);
    const double *dads = rt.in(_dads, n), *tem = rt.in(_tem, n), *item = rt.in(_item, n);
    const double *iqx = rt.in(_iqx, n), *iqy = rt.in(_iqy, n), *iqz = rt.in(_iqz, n);
    const double *g = rt.in(_g, ng, Q<T>::nc), *ig = rt.in(_ig, ng, Q<T>::nc);
    const double *diff = rt.in(_diffusivity, n), *dkds = rt.in(_dkds, n);
    double *dfds = rt.inout(_dfds, n);
    gpu::check(lbm_sensitivity_brinkman_diffusivity(_q.handle(), dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx,
                                                    iqy, iqz, g, ig, diff, dkds),
               "AAD::SensitivityBrinkmanDiffusivity");
    rt.wrote({_dfds}, true);
}
template <class T, template <class> class Q, class Fv, class Ff>
void SensitivityTemperatureAtHeatSource(Q<T> &_q, T *_dfds, const T *_ux, const T *_uy, const T *_imx,
                                        const T *_imy, const T *_dads, const T *_tem, const T *_item, const T *_iqx,
                                        const T *_iqy, const T *_g, const T *_ig, const T *_diffusivity,
                                        const T *_dkds, Fv _qnbc, Ff _bctype) {  // :1497-1517
    static_assert(Q<T>::nd == 2, "2D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    size_t n = gpu::N(_q), ng = n * Q<T>::nc;
    std::vector<int> m = _q.faces_int(_bctype);
    std::vector<double> qn = _q.faces_val(_qnbc, m);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *imx = rt.in(_imx, n), *imy = rt.in(_imy, n);
    const double *dads = rt.in(_dads, n), *tem = rt.in(_tem, n), *item = rt.in(_item, n);
    const double *iqx = rt.in(_iqx, n), *iqy = rt.in(_iqy, n);
    const double *g = rt.in(_g, ng, Q<T>::nc), *ig = rt.in(_ig, ng, Q<T>::nc);
    const double *diff = rt.in(_diffusivity, n), *dkds = rt.in(_dkds, n);
    double *dfds = rt.inout(_dfds, n);
    gpu::check(lbm_sensitivity_temperature_at_heat_source(_q.handle(), dfds, ux, uy, nullptr, imx, imy, nullptr, dads,
                                                          tem, item, iqx, iqy, nullptr, g, ig, diff, dkds, m.data(),
                                                          qn.data()),
               "AAD::SensitivityTemperatureAtHeatSource");
    rt.wrote({_dfds}, true);
}
// :1519-1543 with the face terms of :805-888 taken in their declared argument order (the plain
// reference passes them out of order; its AVX overloads do not)
template <class T, template <class> class Q, class Fv, class Ff>
void SensitivityTemperatureAtHeatSource(Q<T> &_q, T *_dfds, const T *_ux, const T *_uy, const T *_uz,
                                        const T *_imx, const T *_imy, const T *_imz, const T *_dads, const T *_tem,
                                        const T *_item, const T *_iqx, const T *_iqy, const T *_iqz, const T *_g,
                                        const T *_ig, const T *_diffusivity, const T *_dkds, Fv _qnbc,
                                        Ff _bctype) {
    static_assert(Q<T>::nd == 3, "3D overload");
    gpu::Runtime &rt = gpu::Runtime::get();
    rt.flush();
    size_t n = gpu::N(_q), ng = n * Q<T>::nc;
    std::vector<int> m = _q.faces_int(_bctype);
    std::vector<double> qn = _q.faces_val(_qnbc, m);
    const double *ux = rt.in(_ux, n), *uy = rt.in(_uy, n), *uz = rt.in(_uz, n);
    const double *imx = rt.in(_imx, n), *imy = rt.in(_imy, n), *imz = rt.in(_imz, n);
    const double *dads = rt.in(_dads, n), *tem = rt.in(_tem, n), *item = rt.in(_item, n);
    const double *iqx = rt.in(_iqx, n), *iqy = rt.in(_iqy, n), *iqz = rt.in(_iqz, n);
    const double *g = rt.in(_g, ng, Q<T>::nc), *ig = rt.in(_ig, ng, Q<T>::nc);
    const double *diff = rt.in(_diffusivity, n), *dkds = rt.in(_dkds, n);
    double *dfds = rt.inout(_dfds, n);
    gpu::check(lbm_sensitivity_temperature_at_heat_source(_q.handle(), dfds, ux, uy, uz, imx, imy, imz, dads, tem, item,
                                                          iqx, iqy, iqz, g, ig, diff, dkds, m.data(), qn.data()),
               "AAD::SensitivityTemperatureAtHeatSource");
    rt.wrote({_dfds}, true);
}
}  // namespace AAD

}  // namespace PANSLBM2

#include "panslbm2_amd_vtk.hpp"  // VTKExport, VTKXMLExport, VTKImport (host I/O)
