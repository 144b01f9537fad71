// oldefine.h -- drop-in for the reference's include/oldefine.h (enums kept identical
// in names and values so existing callers compile unchanged).
#ifndef OLDEFINE_H
#define OLDEFINE_H

// single-precision pi spelling used by the reference's GPU kernels
// (include/oldefine.h:4); the MI355X kernels reduce phases in FP64 instead
#define F32_PI 3.14159265f

namespace ol {
// include/oldefine.h:7-9
enum PROPMODE { ASM, FRESNEL, AUTO };
// include/oldefine.h:11-13
enum HOLOGRAMSTEP { RECORD, PLAY };
}  // namespace ol
#endif
