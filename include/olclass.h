// olclass.h -- point-cloud value types with the reference's interface
// (include/olclass.h:44-70): ol::Vect3<T> {x, y, z} and ol::Object<T> {size, points, operator[]}.
#ifndef OLCLASS_H
#define OLCLASS_H
#include <cstddef>
#include <memory>

namespace ol {

template <typename _Tp>
struct Vect3 {
    _Tp x, y, z;
    Vect3(_Tp x_, _Tp y_, _Tp z_) : x(x_), y(y_), z(z_) {}
    Vect3() {}
};

template <typename _Tp>
struct Object {
    int size = 0;
    std::unique_ptr<Vect3<_Tp>[]> points;
    Object() {}
    explicit Object(int N) : size(N), points(std::make_unique<Vect3<_Tp>[]>(N)) {}
    Vect3<_Tp>& operator[](int n) { return points[n]; }
    const Vect3<_Tp>& operator[](int n) const { return points[n]; }
};

}  // namespace ol
#endif
