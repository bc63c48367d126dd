/* Umbrella header for libtprims (MI355X build). */
#ifndef TPRIMS_H
#define TPRIMS_H
#include "tprims/core.h"
#include "tprims/tapp_ext.h"
#include "tprims/hip.h"
#endif
