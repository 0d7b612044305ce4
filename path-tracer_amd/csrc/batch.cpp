// batch.cpp - batching of many scenes and the serving loop (filled in below).
#include "pt_runtime.h"
