// fl_internal.h -- shared declarations of the library's translation units.
#pragma once
#include "../../include/fastlem_hip.h"

struct fl_rng;
