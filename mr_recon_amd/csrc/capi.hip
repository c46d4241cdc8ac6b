// Library-wide C-ABI helpers: error text and version.
#include <string.h>
#include "capi_util.h"

namespace mra {
static thread_local char g_last_error[512] = "";
void set_last_error(const char* msg) {
  strncpy(g_last_error, msg, sizeof(g_last_error) - 1);
  g_last_error[sizeof(g_last_error) - 1] = '\0';
}
}  // namespace mra

extern "C" {

const char* mra_last_error() { return mra::g_last_error; }

int mra_version() { return 100; }  // 1.0.0

}  // extern "C"
