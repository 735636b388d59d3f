// Serving loop: continuous batching of sweep requests on one GPU.
#pragma once
#include "elp_internal.h"
