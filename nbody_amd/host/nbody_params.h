// nbody_params.h -- string-keyed parameter map for the engine and solver
// factories.  Plays the role of the QVariantMap the reference factories read
// (nbody/nbody_engines.cpp:4-193, nbody/nbody_solvers.cpp:17-110): every key is
// optional and falls back to the factory default; a value that does not parse
// as the requested type makes the getter report failure.
#pragma once

#include <cstdlib>
#include <map>
#include <string>

class nbody_params
{
	std::map<std::string, std::string> m_values;
public:
	nbody_params() {}

	// Parse "key=value" pairs separated by ';' (not ',', which device lists use).
	static nbody_params parse(const std::string& text)
	{
		nbody_params p;
		size_t pos = 0;
		while(pos <= text.size())
		{
			size_t end = text.find(';', pos);
			if(end == std::string::npos) { end = text.size(); }
			std::string item = text.substr(pos, end - pos);
			size_t eq = item.find('=');
			if(eq != std::string::npos)
			{
				std::string k = trim(item.substr(0, eq));
				std::string v = trim(item.substr(eq + 1));
				if(!k.empty()) { p.m_values[k] = v; }
			}
			pos = end + 1;
		}
		return p;
	}

	void set(const std::string& key, const std::string& value) { m_values[key] = value; }
	// "key=value" pairs joined by ';' in key order (the text parse() reads).
	std::string str() const
	{
		std::string out;
		for(const auto& kv : m_values)
		{
			if(!out.empty()) { out += ';'; }
			out += kv.first + '=' + kv.second;
		}
		return out;
	}
	bool contains(const std::string& key) const { return m_values.count(key) != 0; }

	std::string value(const std::string& key, const std::string& def) const
	{
		auto it = m_values.find(key);
		return it == m_values.end() ? def : it->second;
	}

	// Numeric getters: return false when the key is present but malformed.
	bool get_double(const std::string& key, double def, double* out) const
	{
		auto it = m_values.find(key);
		if(it == m_values.end()) { *out = def; return true; }
		const char* s = it->second.c_str();
		char* e = nullptr;
		double v = std::strtod(s, &e);
		if(e == s || *e != '\0') { return false; }
		*out = v;
		return true;
	}

	bool get_long(const std::string& key, long long def, long long* out) const
	{
		auto it = m_values.find(key);
		if(it == m_values.end()) { *out = def; return true; }
		const char* s = it->second.c_str();
		char* e = nullptr;
		long long v = std::strtoll(s, &e, 10);
		if(e == s || *e != '\0') { return false; }
		*out = v;
		return true;
	}

	// "true"/"false"/"1"/"0" as QVariant::toBool accepts them.
	bool get_bool(const std::string& key, bool def, bool* out) const
	{
		auto it = m_values.find(key);
		if(it == m_values.end()) { *out = def; return true; }
		const std::string& s = it->second;
		if(s == "true" || s == "1") { *out = true; return true; }
		if(s == "false" || s == "0" || s.empty()) { *out = false; return true; }
		return false;
	}

private:
	static std::string trim(const std::string& s)
	{
		size_t b = s.find_first_not_of(" \t\n\r");
		if(b == std::string::npos) { return std::string(); }
		size_t e = s.find_last_not_of(" \t\n\r");
		return s.substr(b, e - b + 1);
	}
};
