// Process-wide log sink (source-compatible with libumbreon's log.hpp).
// Default sink: Warning lines to stderr, Info dropped.
#pragma once

#include <functional>

namespace umbreon {

enum class LogLevel { Warning, Info };

using LogSink = std::function<void(LogLevel level, const char* text)>;

// Install a sink (an empty function restores the default).
void setLogSink(LogSink sink);

void logMessage(LogLevel level, const char* fmt, ...)
#if defined(__GNUC__) || defined(__clang__)
    __attribute__((format(printf, 2, 3)))
#endif
    ;

}  // namespace umbreon
