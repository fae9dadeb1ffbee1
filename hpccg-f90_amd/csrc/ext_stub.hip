// Release hook for buffers owned by the optional paths (mixed precision, batched, learned).
#include "context.hpp"

extern "C" void hpccg_free_ext(hpccg_ctx *c)
{
    for (void *&p : c->ext) {
        if (p) hipFree(p);
        p = nullptr;
    }
}
