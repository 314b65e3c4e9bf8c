// Umbrella for the renderer-agnostic public types.
#pragma once

#include "render/edge_types.hpp"
#include "render/frame_result.hpp"
#include "render/render_options.hpp"
#include "render/render_progress.hpp"
#include "log.hpp"
