// Umbrella for the renderer-agnostic public types.
#pragma once

#include "edge_types.hpp"
#include "frame_result.hpp"
#include "render_options.hpp"
#include "render_progress.hpp"
#include "../log.hpp"
