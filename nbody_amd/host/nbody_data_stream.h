// nbody_data_stream.h -- frame files of a run, for dumps and resume.
//
// The reference's stream format (nbody/nbody_data_stream.cpp:60-229,
// nbody/nbody_data_stream_reader.cpp:118-352), written and read with the
// standard library instead of Qt, so that files of either implementation open
// in the other:
//   <base>.idx   text: "#coord_size 8", "#body_count N", "#box_size B", then one
//                line per frame "step<TAB>time<TAB>file_n<TAB>byte offset";
//                times are written with 6 significant digits, as QTextStream
//                writes a double by default;
//   <base>K.dat  frames: N positions then N velocities, AoS binary64 (48 N
//                bytes); a new K starts when the current file has reached
//                max_part_size bytes;
//   <base>.col   N colours, RGBA float (written with the first frame);
//   <base>.mass  N masses, binary64 (written with the first frame).
// Every call returns 0 on success and -1 (with a log line) on failure.
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "nbtype.h"

class nbody_data;
class nbody_data_stream_reader;

class nbody_data_stream
{
	size_t      m_file_n;
	int64_t     m_max_part_size;
	FILE*       m_data;
	FILE*       m_idx;
	std::string m_base_name;
	bool        m_header_written;

	int open_data_file(bool append);
	int open_index_file(bool append);
	int write_header(size_t count, size_t box_size, const nbcolor_t* color, const nbcoord_t* mass);
public:
	nbody_data_stream();
	~nbody_data_stream();
	nbody_data_stream(const nbody_data_stream&) = delete;
	nbody_data_stream& operator=(const nbody_data_stream&) = delete;

	//! Append the current frame of `data` (nbody_data_stream::write, :100-153).
	int write(const nbody_data* data);
	//! The same for a frame held outside an nbody_data: `data` supplies the
	//! header (count, box size, colours, masses), pos/vel the frame (AoS, count
	//! entries each).
	int write_frame(const nbody_data* data, size_t step, nbcoord_t time, const nbvertex_t* pos,
	                const nbvertex_t* vel);
	//! Create the directories of `name` and open <name>0.dat and <name>.idx;
	//! with append_to, continue its last data file and the index instead
	//! (nbody_data_stream::open, :155-190).
	int open(const std::string& name, int64_t max_part_size, const nbody_data_stream_reader* append_to = nullptr);
	void close();

	static std::string make_idx_name(const std::string& file_base_name);
	static std::string make_dat_name(const std::string& file_base_name, size_t file_n);
	static std::string make_col_name(const std::string& file_base_name);
	static std::string make_mass_name(const std::string& file_base_name);
	static char get_idx_separator();
};

class nbody_data_stream_reader
{
	struct item
	{
		size_t    frame;
		size_t    step;
		nbcoord_t time;
		size_t    file_n;
		int64_t   file_pos;
	};
	std::vector<item>           m_frames;
	std::map<size_t, size_t>    m_step2frame;
	std::map<nbcoord_t, size_t> m_time2frame;
	std::string                 m_file_base_name;
	FILE*                       m_file;
	size_t                      m_file_n;
	size_t                      m_current_frame;
	size_t                      m_coord_size;
	size_t                      m_body_count;
	size_t                      m_box_size;
	bool                        m_colors_read;
	bool                        m_masses_read;

	void parse_header_line(const std::string& line);
	int read_colors(nbody_data* data);
	int read_masses(nbody_data* data);
public:
	nbody_data_stream_reader();
	~nbody_data_stream_reader();
	nbody_data_stream_reader(const nbody_data_stream_reader&) = delete;
	nbody_data_stream_reader& operator=(const nbody_data_stream_reader&) = delete;

	//! Parse <base>.idx and seek to frame 0 (nbody_data_stream_reader::load).
	int load(const std::string& file_base_name);
	void close();
	size_t get_frame_count() const;
	size_t get_steps_count() const;
	nbcoord_t get_max_time() const;
	//! Position on a frame; 0 at once when it is already the current one.
	int seek(size_t frame);
	size_t get_current_frame() const;
	size_t get_body_count() const;
	size_t get_coord_size() const;
	nbcoord_t get_box_size() const;
	size_t get_last_file_n() const;
	//! Read the current frame into `data` (which must hold get_body_count()
	//! bodies), set its time and step, move to the next frame; colours and
	//! masses are read with the first frame.
	int read(nbody_data* data);
};
