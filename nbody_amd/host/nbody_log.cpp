#include "nbody_log.h"

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>

namespace {
std::atomic<size_t> g_count(0);
std::atomic<bool>   g_quiet(false);
std::mutex          g_mutex;
std::string         g_last;
}

void nbody_log(const char* fmt, ...)
{
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	{
		std::lock_guard<std::mutex> lock(g_mutex);
		g_last = buf;
	}
	g_count++;
	if(!g_quiet)
	{
		std::fprintf(stderr, "nbody: %s\n", buf);
	}
}

void nbody_print(const char* fmt, ...)
{
	if(g_quiet)
	{
		return;
	}
	va_list ap;
	va_start(ap, fmt);
	std::vfprintf(stdout, fmt, ap);
	va_end(ap);
	std::fputc('\n', stdout);
	std::fflush(stdout);
}

size_t nbody_log_count()
{
	return g_count;
}

const char* nbody_log_last()
{
	static thread_local std::string copy;
	std::lock_guard<std::mutex> lock(g_mutex);
	copy = g_last;
	return copy.c_str();
}

void nbody_log_set_quiet(bool quiet)
{
	g_quiet = quiet;
}
