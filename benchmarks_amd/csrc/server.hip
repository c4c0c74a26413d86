// Request server (implementation pending; entry points report
// D3H_ERR_UNSUPPORTED until then).
#include "runtime.hpp"

using namespace d3h;

extern "C" {
int d3h_server_create(d3h_context *, const d3h_server_config *, d3h_server **out) {
  if (out) *out = nullptr;
  return fail(D3H_ERR_UNSUPPORTED, "server not built yet");
}
int d3h_server_destroy(d3h_server *) { return D3H_OK; }
int d3h_server_submit(d3h_server *, const d3h_problem *, const void *, int, void *, int, int64_t *) {
  return fail(D3H_ERR_UNSUPPORTED, "server");
}
int d3h_server_poll(d3h_server *, int64_t, int *, double *) { return fail(D3H_ERR_UNSUPPORTED, "server"); }
int d3h_server_wait(d3h_server *, int64_t, double *) { return fail(D3H_ERR_UNSUPPORTED, "server"); }
int d3h_server_stats(d3h_server *, int64_t *, int64_t *, int64_t *, double *) {
  return fail(D3H_ERR_UNSUPPORTED, "server");
}
}
