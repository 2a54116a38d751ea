// Umbrella header of the graph_one C++ API on MI355X (the reference's include/GraphOne/graph_one.h).
#pragma once

#include "GraphOne/type.hpp"
#include "GraphOne/utils.h"
#include "GraphOne/archi.h"
#include "GraphOne/buffer.h"
#include "GraphOne/loader.h"
#include "GraphOne/graph.h"
#include "GraphOne/frontier.h"
#include "GraphOne/spmv.h"
#include "GraphOne/engine.h"
