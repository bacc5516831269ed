// Status codes shared by every C-ABI entry point (defined in the public header).
#pragma once
#include "../../include/voltax_mi355x.h"
