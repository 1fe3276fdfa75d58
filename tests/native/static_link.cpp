// A program linked against the static library (libjxg.a + the HIP runtime),
// as a Rust crate's build.rs links it: the host-side entry points run
// without a GPU, and a context creation fails cleanly when none is present.
#include <cstdio>
#include <vector>

#include "jxg.h"

int main() {
  jxg_frame_params p;
  if (jxg_frame_params_init(2.35f, JXG_SPEED_FAST, 1, &p) != JXG_OK) return 1;
  std::vector<float> blob(JXG_QT_BLOB_FLOATS);
  if (jxg_default_quant_tables(2.35f, JXG_ARBITER_SSIMULACRA2, 1, blob.data()) != JXG_OK) return 2;
  jxg_ctx *ctx = nullptr;
  const int rc = jxg_ctx_create(0, nullptr, &ctx);
  std::printf("scale %.9g table0 %.9g ctx %d\n", p.scale, blob[0], rc);
  if (rc == JXG_OK) jxg_ctx_destroy(ctx);
  return 0;
}
