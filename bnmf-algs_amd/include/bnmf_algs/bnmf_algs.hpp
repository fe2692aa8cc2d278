// bnmf_algs.hpp — umbrella header of the drop-in C++ API (the path of
// src/bnmf_algs.hpp that this library provides: bld_mult, bld_add, bld_appr
// and their helpers).
#pragma once

#include "alloc_model/alloc_model_funcs.hpp"
#include "alloc_model/alloc_model_params.hpp"
#include "bld/bld_add.hpp"
#include "bld/bld_appr.hpp"
#include "bld/bld_fact.hpp"
#include "bld/bld_mult.hpp"
#include "defs.hpp"
#include "util/util.hpp"
