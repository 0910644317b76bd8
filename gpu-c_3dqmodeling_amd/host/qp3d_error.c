/* Last-error message of the qp3d library, one per calling thread. */
#include <stdarg.h>
#include <stdio.h>

#include "qp3d.h"

static _Thread_local char g_err[1024];

const char *qp3d_last_error(void) { return g_err; }

void qp3d_set_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
