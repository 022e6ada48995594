// adjoint.cpp -- differentiable collisionless run (C ABI half).
#include "engine.h"
using namespace pic;
extern "C" {
int pic_adjoint_create(const pic_config*, const double*, const double*, int64_t, int,
                       const int32_t*, const int32_t*, int, int, pic_adjoint**) {
  return fail(PIC_ERR_UNSUPPORTED, "adjoint not built yet");
}
int pic_adjoint_destroy(pic_adjoint*) { return PIC_OK; }
int pic_adjoint_forward(pic_adjoint*, const double*, const void*, const void*, double*) {
  return fail(PIC_ERR_UNSUPPORTED, "adjoint not built yet");
}
int pic_adjoint_backward(pic_adjoint*, double, double*, void*, void*) {
  return fail(PIC_ERR_UNSUPPORTED, "adjoint not built yet");
}
int pic_adjoint_stream(pic_adjoint*, void**) { return fail(PIC_ERR_UNSUPPORTED, "adjoint not built yet"); }
}
