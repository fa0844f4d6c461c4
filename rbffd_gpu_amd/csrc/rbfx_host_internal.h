// Shared declarations for the host-only translation units of librbfx.
#pragma once
#include "rbfx.h"
