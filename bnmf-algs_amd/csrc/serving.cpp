// serving.cpp — placeholder translation unit for the serving loop.
#include "bnmf_internal.hpp"
