/*
 * device.c -- GPU state attached to an AQD2.
 *
 * AQD2 has no spare field (its layout is fixed by ex.dat), so device buffers of
 * a model live in a side table keyed by the AQD2 address and are released by
 * finalize_aqd2.
 */
#include "host_internal.h"

void aqd2_release_device(AQD2 *qd)
{
  (void)qd;
}
