// sx_pipeline.h -- device filter cascade (pipeline.rs:159-189) and the host-side
// domain definition / reporting (pipeline.rs:190-294, search.rs:131-271).
#pragma once
#include "sx_runtime.h"
