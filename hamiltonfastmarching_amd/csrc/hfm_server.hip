// hfm_server.hip -- continuous-batching solve server (filled in by the serving driver).
#include "hfm_internal.h"
