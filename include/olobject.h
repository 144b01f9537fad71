// olobject.h -- drop-in for the reference's include/olobject.h: .3d / .3df point-cloud
// readers and the cloud helpers objinfo, objsort and objset, over libolx.so.
//
// Reading is host I/O (olx_obj_read).  objinfo / objsort / objset run on the device
// (olx_obj_bounds / olx_obj_sort_z / olx_obj_set): host Object<T> overloads stage the
// cloud to HBM and back; gObject<T> overloads work in place in HBM.  Object<float>
// clouds get the reference's float arithmetic (OLX_OBJSET_F32), Object<double> its
// double arithmetic, so results match the reference's bit for bit.  objsort is
// stable (std::sort leaves the order of equal depths unspecified).
#ifndef OLOBJECT_H
#define OLOBJECT_H
#include <cstdint>
#include <iostream>
#include <type_traits>
#include <vector>

#include "olclass.h"
#include "oldevice.h"
#include "olx.h"

namespace ol {

// include/olobject.h:193-195
enum OBJSET_MODE { ZMIN, ZMIDDLE };

namespace detail {

inline std::vector<olx_point> read_points(const char* path, int format) {
    int64_t n = 0;
    check(olx_obj_read(path, format, nullptr, 0, &n, 1.0, 0.0, 0.0, 0.0), "objread");
    std::vector<olx_point> v((std::size_t)n);
    if (n) check(olx_obj_read(path, format, v.data(), n, &n, 1.0, 0.0, 0.0, 0.0), "objread");
    return v;
}

template <typename _Tp>
Object<_Tp> to_object(const std::vector<olx_point>& v) {
    Object<_Tp> obj((int)v.size());
    for (std::size_t j = 0; j < v.size(); j++)
        obj[(int)j] = Vect3<_Tp>((_Tp)v[j].x, (_Tp)v[j].y, (_Tp)v[j].z);
    return obj;
}

template <typename _Tp>
void from_points(Object<_Tp>& obj, const std::vector<olx_point>& v) {
    for (int j = 0; j < obj.size; j++) obj[j] = Vect3<_Tp>((_Tp)v[j].x, (_Tp)v[j].y, (_Tp)v[j].z);
}

// a host cloud in a device staging buffer (slot 2), as olx_point records
template <typename _Tp>
olx_point* stage_cloud(const Object<_Tp>& obj) {
    auto v = to_points(obj);
    void* d = staging(2).get(sizeof(olx_point) * (v.empty() ? 1 : v.size()));
    if (!v.empty()) check(olx_copy_h2d(ctx(), d, v.data(), sizeof(olx_point) * v.size()), "stage cloud");
    return static_cast<olx_point*>(d);
}

template <typename _Tp>
void unstage_cloud(Object<_Tp>& obj, const olx_point* d) {
    std::vector<olx_point> v((std::size_t)obj.size);
    if (!v.empty()) check(olx_copy_d2h(ctx(), v.data(), d, sizeof(olx_point) * v.size()), "unstage cloud");
    from_points(obj, v);
}

template <typename _Tp>
int objset_mode(OBJSET_MODE mode) {
    return (mode == ZMIDDLE ? OLX_OBJSET_ZMIDDLE : OLX_OBJSET_ZMIN) |
           (std::is_same<_Tp, float>::value ? OLX_OBJSET_F32 : 0);
}

// the three lines objinfo prints (include/olobject.h:171-173), values in _Tp
template <typename _Tp>
void print_bounds(const olx_point* d, int n, bool ranges) {
    double mn[3], mx[3];
    check(olx_obj_bounds(ctx(), d, n, mn, mx), "objinfo");
    const char* ax[3] = {"x", "y", "z"};
    for (int c = 0; c < 3; c++) {
        std::cout << ax[c] << "max: " << (_Tp)mx[c] << "  " << ax[c] << "min: " << (_Tp)mn[c];
        if (ranges) std::cout << "  range" << (_Tp)((_Tp)mx[c] - (_Tp)mn[c]);
        std::cout << std::endl;
    }
}

}  // namespace detail

// ------------------------------------------------------------------ readers
// include/olobject.h:9-42
template <typename _Tp>
Object<_Tp> objread_3d(const char* path) {
    return detail::to_object<_Tp>(detail::read_points(path, OLX_OBJ_3D));
}
// include/olobject.h:44-77
template <typename _Tp>
Object<_Tp> objread_3df(const char* path) {
    return detail::to_object<_Tp>(detail::read_points(path, OLX_OBJ_3DF));
}
// include/olobject.h:79-91 (an unknown extension throws instead of exiting)
template <typename _Tp>
Object<_Tp> objread(const char* path) {
    return detail::to_object<_Tp>(detail::read_points(path, OLX_OBJ_AUTO));
}
// include/olobject.h:93-126: int32 triples, x = x_int * scale + offset[0] in the
// caller's types (the integers are read exactly, the affine map is the reference's
// expression)
template <typename _Tp, typename OFFSET_T, typename SCALE_T>
Object<_Tp> objread(const char* path, OFFSET_T* offset, SCALE_T scale) {
    auto v = detail::read_points(path, OLX_OBJ_3D);
    Object<_Tp> obj((int)v.size());
    for (int j = 0; j < obj.size; j++) {
        const int32_t xi = (int32_t)v[j].x, yi = (int32_t)v[j].y, zi = (int32_t)v[j].z;
        obj[j].x = xi * scale + offset[0];
        obj[j].y = yi * scale + offset[1];
        obj[j].z = zi * scale + offset[2];
    }
    return obj;
}
// include/olobject.h:128-161
template <typename _Tp, typename OFFSET_T, typename SCALE_T>
Object<_Tp> objread(const char* path, Vect3<OFFSET_T> offset, SCALE_T scale) {
    OFFSET_T o[3] = {offset.x, offset.y, offset.z};
    return objread<_Tp>(path, o, scale);
}

// ------------------------------------------------------------------ helpers
// include/olobject.h:163-174: prints the bounding box
template <typename _Tp>
void objinfo(Object<_Tp>& obj) {
    detail::print_bounds<_Tp>(detail::stage_cloud(obj), obj.size, false);
}
template <typename _Tp>
void objinfo(gObject<_Tp>& obj) {
    detail::print_bounds<_Tp>(obj.points.get(), obj.size, false);
}

// include/olobject.h:176-181: descending z (stable)
template <typename _Tp>
void objsort(Object<_Tp>& obj) {
    olx_point* d = detail::stage_cloud(obj);
    detail::check(olx_obj_sort_z(detail::ctx(), d, obj.size), "objsort");
    detail::unstage_cloud(obj, d);
}
template <typename _Tp>
void objsort(gObject<_Tp>& obj) {
    detail::check(olx_obj_sort_z(detail::ctx(), obj.points.get(), obj.size), "objsort");
}

// include/olobject.h:196-228: the larger of the x / y extents fits xy_range (prints
// the bounds and ranges as the reference does)
template <typename _Tp>
void objset(Object<_Tp>& obj, float xy_range, Vect3<_Tp> offset, OBJSET_MODE mode = ZMIN) {
    olx_point* d = detail::stage_cloud(obj);
    detail::print_bounds<_Tp>(d, obj.size, true);
    detail::check(olx_obj_set(detail::ctx(), d, obj.size, xy_range, -1.0, offset.x, offset.y,
                              offset.z, detail::objset_mode<_Tp>(mode)), "objset");
    detail::unstage_cloud(obj, d);
}
// include/olobject.h:232-264: the smaller of x_range / range_x and y_range / range_y
template <typename _Tp>
void objset(Object<_Tp>& obj, float x_range, float y_range, Vect3<_Tp> offset,
            OBJSET_MODE mode = ZMIN) {
    olx_point* d = detail::stage_cloud(obj);
    detail::check(olx_obj_set(detail::ctx(), d, obj.size, x_range, y_range, offset.x, offset.y,
                              offset.z, detail::objset_mode<_Tp>(mode)), "objset");
    detail::unstage_cloud(obj, d);
}
// device-resident forms (no reference counterpart: the reference fits clouds on the host)
template <typename _Tp>
void objset(gObject<_Tp>& obj, float x_range, float y_range, Vect3<_Tp> offset,
            OBJSET_MODE mode = ZMIN) {
    detail::check(olx_obj_set(detail::ctx(), obj.points.get(), obj.size, x_range, y_range,
                              offset.x, offset.y, offset.z, detail::objset_mode<_Tp>(mode)),
                  "objset");
}

}  // namespace ol
#endif
