// nbody_log.h -- diagnostics of the host layer.
//
// The reference reports interface misuse with qDebug() and returns
// (e.g. nbody/nbody_engine_simple.cpp:122-131).  Here messages go to stderr
// (unless silenced) and are counted, so a caller can check that a call was
// rejected without parsing text.
#pragma once

#include <cstddef>

void nbody_log(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
// Plain informational line on stdout (statistics, engine info).
void nbody_print(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
size_t nbody_log_count();
const char* nbody_log_last();
void nbody_log_set_quiet(bool quiet);
