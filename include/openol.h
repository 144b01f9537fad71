// openol.h -- umbrella header of the MI355X drop-in (role of the reference's
// include/openol.h:3-22): the propagation, CGH, array-helper, point-cloud and image
// APIs over libolx.so.
// Link with -L<repo>/openol_amd -lolx (rpath to the same directory).
#ifndef OPENOL_H
#define OPENOL_H
#include <complex>
#include <memory>

#include "Diffuser.h"
#include "Lensarry.h"
#include "olcgh.h"
#include "olclass.h"
#include "oldefine.h"
#include "oldevice.h"
#include "olimg.h"
#include "olobject.h"
#include "olprop.h"
#include "olutils.h"
#include "olx.h"
#endif
