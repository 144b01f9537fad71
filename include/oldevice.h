// oldevice.h -- device-resident buffers and point clouds for the drop-in headers.
//
// The reference's device API uses cuda::unique_ptr<thrust::complex<T>[]>,
// cuda::make_unique, cpu2cuda / cuda2cpu and gObject (include/golclass.h:23-121).
// Here the same roles are ol::device_ptr<T[]>, ol::make_device<T[]>, cpu2dev /
// dev2cpu and ol::gObject, all backed by libolx.so (include/olx.h); complex
// device fields are std::complex<float> with the interleaved (re, im) layout.
// Errors throw std::runtime_error with the library's message (the reference
// throws only from its allocator and otherwise exits the process).
#ifndef OLDEVICE_H
#define OLDEVICE_H
#include <complex>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "olclass.h"
#include "olx.h"

namespace ol {
namespace detail {

inline void check(int rc, const char* what) {
    if (rc != OLX_OK) throw std::runtime_error(std::string(what) + ": " + olx_last_error());
}

// one library context per process, on device OLX_DEVICE (default 0), default stream
inline olx_ctx* ctx() {
    static olx_ctx* c = [] {
        olx_ctx* p = nullptr;
        const char* e = std::getenv("OLX_DEVICE");
        check(olx_ctx_create(&p, e ? std::atoi(e) : 0), "olx_ctx_create");
        return p;
    }();
    return c;
}

// grow-only device staging buffers for the host-data overloads
struct Staging {
    void* p = nullptr;
    size_t bytes = 0;
    void* get(size_t b) {
        if (b > bytes) {
            if (p) olx_dev_free(ctx(), p);
            p = nullptr;
            bytes = 0;
            check(olx_dev_alloc(ctx(), &p, b), "staging allocation");
            bytes = b;
        }
        return p;
    }
};
inline Staging& staging(int slot) {
    static Staging s[3];
    return s[slot];
}

}  // namespace detail

// ------------------------------------------------------------------ device_ptr
struct device_deleter {
    void operator()(void* p) const {
        if (p) olx_dev_free(detail::ctx(), p);
    }
};
template <typename T>
using device_ptr = std::unique_ptr<T, device_deleter>;

template <typename T>
typename std::enable_if<std::is_array<T>::value, device_ptr<T>>::type make_device(std::size_t n) {
    using U = typename std::remove_extent<T>::type;
    void* p = nullptr;
    detail::check(olx_dev_alloc(detail::ctx(), &p, sizeof(U) * n), "make_device");
    return device_ptr<T>(static_cast<U*>(p));
}

// host <-> device copies (roles of cpu2cuda / cuda2cpu, include/golclass.h:87-121)
template <typename T>
device_ptr<T[]> cpu2dev(const std::unique_ptr<T[]>& src, std::size_t n) {
    auto d = make_device<T[]>(n);
    detail::check(olx_copy_h2d(detail::ctx(), d.get(), src.get(), sizeof(T) * n), "cpu2dev");
    return d;
}
template <typename T>
void cpu2dev(const std::unique_ptr<T[]>& src, device_ptr<T[]>& dst, std::size_t n) {
    if (!dst) dst = make_device<T[]>(n);
    detail::check(olx_copy_h2d(detail::ctx(), dst.get(), src.get(), sizeof(T) * n), "cpu2dev");
}
template <typename T>
std::unique_ptr<T[]> dev2cpu(const device_ptr<T[]>& src, std::size_t n) {
    auto h = std::make_unique<T[]>(n);
    detail::check(olx_copy_d2h(detail::ctx(), h.get(), src.get(), sizeof(T) * n), "dev2cpu");
    return h;
}
template <typename T>
void dev2cpu(const device_ptr<T[]>& src, std::unique_ptr<T[]>& dst, std::size_t n) {
    if (!dst) dst = std::make_unique<T[]>(n);
    detail::check(olx_copy_d2h(detail::ctx(), dst.get(), src.get(), sizeof(T) * n), "dev2cpu");
}

// ------------------------------------------------------------------ point clouds
// Device point cloud (role of gObject, include/golclass.h:58-71) holding olx_point
// records: x, y, z in double plus per-point amplitude and phase (1 and 0 unless set).
template <typename _Tp>
struct gObject {
    int size = 0;
    device_ptr<olx_point[]> points;
    gObject() {}
};

template <typename _Tp>
std::vector<olx_point> to_points(const Object<_Tp>& obj, const float* amp = nullptr,
                                 const float* phase = nullptr) {
    std::vector<olx_point> v((std::size_t)obj.size);
    for (int j = 0; j < obj.size; j++) {
        v[j].x = (double)obj.points[j].x;
        v[j].y = (double)obj.points[j].y;
        v[j].z = (double)obj.points[j].z;
        v[j].amp = amp ? amp[j] : 1.0f;
        v[j].phase = phase ? phase[j] : 0.0f;
    }
    return v;
}

// role of Object2cuda (include/golclass.h:73-78); amp / phase optional per point
template <typename _Tp>
gObject<_Tp> Object2dev(const Object<_Tp>& obj, const float* amp = nullptr,
                        const float* phase = nullptr) {
    gObject<_Tp> g;
    g.size = obj.size;
    auto v = to_points(obj, amp, phase);
    g.points = make_device<olx_point[]>(v.empty() ? 1 : v.size());
    if (!v.empty())
        detail::check(olx_copy_h2d(detail::ctx(), g.points.get(), v.data(),
                                   sizeof(olx_point) * v.size()), "Object2dev");
    return g;
}

// wait for all queued device work (the reference synchronises after every launch)
inline void synchronize() { detail::check(olx_ctx_synchronize(detail::ctx()), "synchronize"); }

}  // namespace ol
#endif
