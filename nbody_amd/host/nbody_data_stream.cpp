// nbody_data_stream.cpp -- see nbody_data_stream.h.  Behaviour follows the
// reference's writer (nbody/nbody_data_stream.cpp:60-229) and reader
// (nbody/nbody_data_stream_reader.cpp:118-352), including which calls fail and
// how a reader continues after a frame it cannot reach.
#include "nbody_data_stream.h"

#include <sys/stat.h>

#include <cerrno>
#include <cinttypes>
#include <cstdlib>
#include <cstring>
#include <limits>

#include "nbody_data.h"
#include "nbody_log.h"

namespace {

// mkdir -p of every directory above `path` (QFileInfo(name).dir().mkpath(".")).
void make_parent_dirs(const std::string& path)
{
	const size_t slash = path.find_last_of('/');
	if(slash == std::string::npos || slash == 0)
	{
		return;
	}
	const std::string dir = path.substr(0, slash);
	for(size_t pos = 1; pos <= dir.size(); ++pos)
	{
		if(pos == dir.size() || dir[pos] == '/')
		{
			(void)::mkdir(dir.substr(0, pos).c_str(), 0777);
		}
	}
}

// A regular file opened with `mode`, or null (a directory of that name fails,
// as QFile::open does).
FILE* open_regular(const std::string& name, const char* mode)
{
	struct stat st;
	if(::stat(name.c_str(), &st) == 0 && S_ISDIR(st.st_mode))
	{
		errno = EISDIR;
		return nullptr;
	}
	return std::fopen(name.c_str(), mode);
}

bool write_all(FILE* f, const void* p, size_t bytes)
{
	return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes;
}

bool parse_u64(const std::string& s, uint64_t* out)
{
	if(s.empty())
	{
		return false;
	}
	char* end = nullptr;
	errno = 0;
	const unsigned long long v = std::strtoull(s.c_str(), &end, 10);
	if(errno != 0 || end == s.c_str() || *end != '\0' || s[0] == '-')
	{
		return false;
	}
	*out = v;
	return true;
}

bool parse_i64(const std::string& s, int64_t* out)
{
	if(s.empty())
	{
		return false;
	}
	char* end = nullptr;
	errno = 0;
	const long long v = std::strtoll(s.c_str(), &end, 10);
	if(errno != 0 || end == s.c_str() || *end != '\0')
	{
		return false;
	}
	*out = v;
	return true;
}

bool parse_double(const std::string& s, double* out)
{
	if(s.empty())
	{
		return false;
	}
	char* end = nullptr;
	const double v = std::strtod(s.c_str(), &end);
	if(end == s.c_str() || *end != '\0')
	{
		return false;
	}
	*out = v;
	return true;
}

std::vector<std::string> split(const std::string& s, char sep)
{
	std::vector<std::string> parts;
	size_t start = 0;
	while(true)
	{
		const size_t p = s.find(sep, start);
		parts.push_back(s.substr(start, p == std::string::npos ? std::string::npos : p - start));
		if(p == std::string::npos)
		{
			break;
		}
		start = p + 1;
	}
	return parts;
}

}  // namespace

// ---------------------------------------------------------------------------
// writer
// ---------------------------------------------------------------------------
nbody_data_stream::nbody_data_stream() :
	m_file_n(0),
	m_max_part_size(1024 * 1024 * 1024),
	m_data(nullptr),
	m_idx(nullptr),
	m_header_written(false)
{
}

nbody_data_stream::~nbody_data_stream()
{
	close();
}

int nbody_data_stream::open_data_file(bool append)
{
	if(m_data != nullptr)
	{
		std::fclose(m_data);
		m_data = nullptr;
	}
	const std::string name = make_dat_name(m_base_name, m_file_n);
	m_data = open_regular(name, append ? "ab" : "wb");
	if(m_data == nullptr)
	{
		nbody_log("stream: can't open file %s: %s", name.c_str(), std::strerror(errno));
		return -1;
	}
	if(append && fseeko(m_data, 0, SEEK_END) != 0)
	{
		nbody_log("stream: can't seek to the end of %s", name.c_str());
		return -1;
	}
	return 0;
}

int nbody_data_stream::open_index_file(bool append)
{
	if(m_idx != nullptr)
	{
		std::fclose(m_idx);
		m_idx = nullptr;
	}
	const std::string name = make_idx_name(m_base_name);
	m_idx = open_regular(name, append ? "a" : "w");
	if(m_idx == nullptr)
	{
		nbody_log("stream: can't open file %s: %s", name.c_str(), std::strerror(errno));
		return -1;
	}
	return 0;
}

int nbody_data_stream::write_header(size_t count, size_t box_size, const nbcolor_t* color, const nbcoord_t* mass)
{
	std::fprintf(m_idx, "#coord_size %zu\n#body_count %zu\n#box_size %zu\n", sizeof(nbcoord_t), count, box_size);
	const std::string col_name = make_col_name(m_base_name);
	FILE* col = open_regular(col_name, "wb");
	if(col == nullptr)
	{
		nbody_log("stream: can't open file %s: %s", col_name.c_str(), std::strerror(errno));
		return -1;
	}
	const bool col_ok = write_all(col, color, sizeof(nbcolor_t) * count);
	std::fclose(col);
	if(!col_ok)
	{
		nbody_log("stream: can't write file %s", col_name.c_str());
		return -1;
	}
	const std::string mass_name = make_mass_name(m_base_name);
	FILE* mf = open_regular(mass_name, "wb");
	if(mf == nullptr)
	{
		nbody_log("stream: can't open file %s: %s", mass_name.c_str(), std::strerror(errno));
		return -1;
	}
	const bool mass_ok = write_all(mf, mass, sizeof(nbcoord_t) * count);
	std::fclose(mf);
	if(!mass_ok)
	{
		nbody_log("stream: can't write file %s", mass_name.c_str());
		return -1;
	}
	m_header_written = true;
	return 0;
}

int nbody_data_stream::write(const nbody_data* data)
{
	if(data == nullptr)
	{
		nbody_log("stream: data == NULL");
		return -1;
	}
	return write_frame(data, data->get_step(), data->get_time(), data->get_vertites(), data->get_velosites());
}

int nbody_data_stream::write_frame(const nbody_data* data, size_t step, nbcoord_t time, const nbvertex_t* pos,
                                   const nbvertex_t* vel)
{
	if(data == nullptr)
	{
		nbody_log("stream: data == NULL");
		return -1;
	}
	if(m_idx == nullptr || m_data == nullptr)
	{
		nbody_log("stream: not open");
		return -1;
	}
	const size_t count = data->get_count();
	if(!m_header_written && write_header(count, data->get_box_size(), data->get_color(), data->get_mass()) != 0)
	{
		nbody_log("stream: writing the header failed");
		return -1;
	}
	if(m_max_part_size > 0 && ftello(m_data) >= m_max_part_size)
	{
		++m_file_n;
		if(open_data_file(false) != 0)
		{
			return -1;
		}
	}
	const int64_t fpos = ftello(m_data);
	const size_t sz = sizeof(nbvertex_t) * count;
	if(!write_all(m_data, pos, sz) || !write_all(m_data, vel, sz) || std::fflush(m_data) != 0)
	{
		nbody_log("stream: can't write file %s", make_dat_name(m_base_name, m_file_n).c_str());
		return -1;
	}
	const char sep = get_idx_separator();
	std::fprintf(m_idx, "%zu%c%g%c%zu%c%" PRId64 "\n", step, sep, time, sep, m_file_n, sep, fpos);
	std::fflush(m_idx);
	return 0;
}

int nbody_data_stream::open(const std::string& name, int64_t max_part_size, const nbody_data_stream_reader* append_to)
{
	make_parent_dirs(name);
	m_base_name = name;
	m_max_part_size = max_part_size;
	const bool append = append_to != nullptr;
	if(append)
	{
		m_file_n = append_to->get_last_file_n();
		m_header_written = true;
	}
	else
	{
		m_file_n = 0;
	}
	if(open_data_file(append) != 0 || open_index_file(append) != 0)
	{
		return -1;
	}
	return 0;
}

void nbody_data_stream::close()
{
	m_base_name.clear();
	m_file_n = 0;
	if(m_idx != nullptr)
	{
		std::fclose(m_idx);
		m_idx = nullptr;
	}
	if(m_data != nullptr)
	{
		std::fclose(m_data);
		m_data = nullptr;
	}
}

std::string nbody_data_stream::make_idx_name(const std::string& file_base_name)
{
	return file_base_name + ".idx";
}

std::string nbody_data_stream::make_dat_name(const std::string& file_base_name, size_t file_n)
{
	return file_base_name + std::to_string(file_n) + ".dat";
}

std::string nbody_data_stream::make_col_name(const std::string& file_base_name)
{
	return file_base_name + ".col";
}

std::string nbody_data_stream::make_mass_name(const std::string& file_base_name)
{
	return file_base_name + ".mass";
}

char nbody_data_stream::get_idx_separator()
{
	return '\t';
}

// ---------------------------------------------------------------------------
// reader
// ---------------------------------------------------------------------------
nbody_data_stream_reader::nbody_data_stream_reader() :
	m_file(nullptr),
	m_file_n(std::numeric_limits<size_t>::max()),
	m_current_frame(std::numeric_limits<size_t>::max()),
	m_coord_size(0),
	m_body_count(0),
	m_box_size(0),
	m_colors_read(false),
	m_masses_read(false)
{
}

nbody_data_stream_reader::~nbody_data_stream_reader()
{
	close();
}

void nbody_data_stream_reader::parse_header_line(const std::string& line)
{
	const std::vector<std::string> parts = split(line, ' ');
	uint64_t value = 0;
	if(parts.size() != 2 || !parse_u64(parts[1], &value))
	{
		nbody_log("stream: invalid header line %s", line.c_str());
		return;
	}
	if(parts[0] == "#coord_size")
	{
		m_coord_size = value;
	}
	else if(parts[0] == "#body_count")
	{
		m_body_count = value;
	}
	else if(parts[0] == "#box_size")
	{
		m_box_size = value;
	}
}

int nbody_data_stream_reader::read_colors(nbody_data* data)
{
	const std::string name = nbody_data_stream::make_col_name(m_file_base_name);
	FILE* f = open_regular(name, "rb");
	if(f == nullptr)
	{
		nbody_log("stream: can't open file %s", name.c_str());
		return -1;
	}
	const size_t sz = sizeof(nbcolor_t) * data->get_count();
	const bool ok = std::fread(data->get_color(), 1, sz, f) == sz;
	std::fclose(f);
	if(!ok)
	{
		nbody_log("stream: can't read file %s", name.c_str());
		return -1;
	}
	return 0;
}

int nbody_data_stream_reader::read_masses(nbody_data* data)
{
	const std::string name = nbody_data_stream::make_mass_name(m_file_base_name);
	FILE* f = open_regular(name, "rb");
	if(f == nullptr)
	{
		nbody_log("stream: can't open file %s", name.c_str());
		return -1;
	}
	const size_t sz = sizeof(nbcoord_t) * data->get_count();
	const bool ok = std::fread(data->get_mass(), 1, sz, f) == sz;
	std::fclose(f);
	if(!ok)
	{
		nbody_log("stream: can't read file %s", name.c_str());
		return -1;
	}
	return 0;
}

int nbody_data_stream_reader::load(const std::string& file_base_name)
{
	close();
	const std::string idx_name = nbody_data_stream::make_idx_name(file_base_name);
	FILE* idx = open_regular(idx_name, "r");
	if(idx == nullptr)
	{
		nbody_log("stream: can't open index file %s", idx_name.c_str());
		return -1;
	}
	std::string line;
	int ch = 0;
	bool bad = false;
	while(!bad)
	{
		line.clear();
		while((ch = std::fgetc(idx)) != EOF && ch != '\n')
		{
			line.push_back(static_cast<char>(ch));
		}
		if(ch == EOF && line.empty())
		{
			break;
		}
		if(!line.empty() && line[0] == '#')
		{
			parse_header_line(line);
			continue;
		}
		const std::vector<std::string> parts = split(line, nbody_data_stream::get_idx_separator());
		if(parts.size() != 4)
		{
			nbody_log("stream: incomplete data line %s", line.c_str());
			bad = true;
			break;
		}
		item it;
		uint64_t step = 0, file_n = 0;
		it.frame = m_frames.size();
		if(!parse_u64(parts[0], &step) || !parse_double(parts[1], &it.time) || !parse_u64(parts[2], &file_n) ||
		   !parse_i64(parts[3], &it.file_pos))
		{
			nbody_log("stream: invalid data line %s", line.c_str());
			bad = true;
			break;
		}
		it.step = step;
		it.file_n = file_n;
		m_frames.push_back(it);
		m_step2frame[it.step] = it.frame;
		m_time2frame[it.time] = it.frame;
	}
	std::fclose(idx);
	if(bad)
	{
		return -1;
	}
	if(m_body_count == 0 || m_coord_size == 0 || m_box_size == 0)
	{
		nbody_log("stream: invalid file header: coord_size %zu body_count %zu box_size %zu", m_coord_size,
		          m_body_count, m_box_size);
		return -1;
	}
	if(m_coord_size != sizeof(nbcoord_t))
	{
		nbody_log("stream: invalid file header: coord_size %zu must be %zu", m_coord_size, sizeof(nbcoord_t));
		return -1;
	}
	m_file_base_name = file_base_name;
	return seek(0);
}

void nbody_data_stream_reader::close()
{
	m_frames.clear();
	m_step2frame.clear();
	m_time2frame.clear();
	m_file_base_name.clear();
	if(m_file != nullptr)
	{
		std::fclose(m_file);
		m_file = nullptr;
	}
	m_file_n = std::numeric_limits<size_t>::max();
	m_current_frame = std::numeric_limits<size_t>::max();
}

size_t nbody_data_stream_reader::get_frame_count() const
{
	return m_frames.size();
}

size_t nbody_data_stream_reader::get_steps_count() const
{
	return m_frames.empty() ? 0 : m_frames.back().step;
}

nbcoord_t nbody_data_stream_reader::get_max_time() const
{
	return m_frames.empty() ? 0 : m_frames.back().time;
}

int nbody_data_stream_reader::seek(size_t frame)
{
	if(frame == m_current_frame)
	{
		return 0;
	}
	if(frame >= m_frames.size())
	{
		nbody_log("stream: can't seek to frame %zu: out of range", frame);
		return -1;
	}
	const item& it = m_frames[frame];
	if(m_file_n != it.file_n || m_file == nullptr)
	{
		if(m_file != nullptr)
		{
			std::fclose(m_file);
			m_file = nullptr;
		}
		const std::string name = nbody_data_stream::make_dat_name(m_file_base_name, it.file_n);
		m_file = open_regular(name, "rb");
		if(m_file == nullptr)
		{
			nbody_log("stream: can't open file %s", name.c_str());
			return -1;
		}
		m_file_n = it.file_n;
	}
	if(it.file_pos < 0 || fseeko(m_file, it.file_pos, SEEK_SET) != 0)
	{
		nbody_log("stream: can't seek file to offset %" PRId64, it.file_pos);
		std::fclose(m_file);
		m_file = nullptr;
		return -1;
	}
	m_current_frame = frame;
	return 0;
}

size_t nbody_data_stream_reader::get_current_frame() const
{
	return m_current_frame;
}

size_t nbody_data_stream_reader::get_body_count() const
{
	return m_body_count;
}

size_t nbody_data_stream_reader::get_coord_size() const
{
	return m_coord_size;
}

nbcoord_t nbody_data_stream_reader::get_box_size() const
{
	return static_cast<nbcoord_t>(m_box_size);
}

size_t nbody_data_stream_reader::get_last_file_n() const
{
	return m_frames.empty() ? 0 : m_frames.back().file_n;
}

int nbody_data_stream_reader::read(nbody_data* data)
{
	if(data == nullptr)
	{
		nbody_log("stream: data == NULL");
		return -1;
	}
	if(data->get_count() != m_body_count)
	{
		nbody_log("stream: destination holds %zu bodies, the stream %zu", data->get_count(), m_body_count);
		return -1;
	}
	if(seek(m_current_frame) != 0)
	{
		nbody_log("stream: can't seek to the current frame %zu", m_current_frame);
		return -1;
	}
	const size_t sz = sizeof(nbvertex_t) * data->get_count();
	if(m_file == nullptr || std::fread(data->get_vertites(), 1, sz, m_file) != sz ||
	   std::fread(data->get_velosites(), 1, sz, m_file) != sz)
	{
		nbody_log("stream: can't read frame %zu", m_current_frame);
		return -1;
	}
	const item& it = m_frames[m_current_frame];
	data->set_time(it.time);
	data->set_step(it.step);
	if(m_current_frame + 1 < m_frames.size() && seek(m_current_frame + 1) != 0)
	{
		nbody_log("stream: seek to frame %zu failed; frame %zu was read", m_current_frame + 1, m_current_frame);
	}
	if(!m_colors_read)
	{
		m_colors_read = read_colors(data) == 0;
	}
	if(!m_masses_read)
	{
		m_masses_read = read_masses(data) == 0;
	}
	return 0;
}
