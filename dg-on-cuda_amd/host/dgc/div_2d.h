// div_2d (src/div_2d.h:36-59) lives with grad_2d: they share the device
// discretization and the weak-form operators.
#ifndef DGC_DIV_2D_H
#define DGC_DIV_2D_H
#include "grad_2d.h"
#endif
