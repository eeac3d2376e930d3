// Restart files in the reference's on-disk formats.
//
// A block's restart payload (".part") is a TorchScript archive of a module
// whose buffers are the variables, by name: the format kintera::save_tensors /
// load_tensors write and read (kintera is an unvendored dependency of the
// reference; python/write_init.py:12-22 writes the same archive from Python
// and python/api/pd_inspect.py:21-60 reads it back through torch.jit.load and
// named_buffers).  Blocks of one output are combined into one ".restart" file:
// a single block's part is renamed, several are concatenated behind a text
// index (src/output/combine_restart.cpp:35-75):
//
//   SNAPY_RESTART_BUNDLE_V1\n
//   <count>\n
//   <part file name>\t<size in bytes>\n      (count lines)
//   \n
//   <payloads, back to back, in index order>
#pragma once

#include <string>
#include <vector>

#include "snap_modules.hpp"

namespace snap {

// kintera::save_tensors: every defined tensor of vars becomes a buffer of one
// TorchScript module, saved to filename
void save_tensors(Variables const& vars, std::string const& filename);

// kintera::load_tensors: the buffers of a TorchScript archive, by name (host
// tensors)
Variables load_tensors(std::string const& filename);

// src/output/combine_restart.cpp:35-75: 0 on success, -1 when a file cannot
// be read or the bundle cannot be written
int make_restart_bundle(std::string const& bundle_name,
                        std::vector<std::string> const& file_list);

// src/input/read_restart_file.cpp:223-235: a bundle (first line is the magic
// string) yields the payload of block<block_rank>; any other file is read as a
// single payload.  A bundle that cannot be parsed, or that holds no payload
// for the block, yields an empty map and a message on stderr, as in the
// reference; a missing or malformed single payload throws.
Variables load_restart(std::string const& path, int block_rank = 0);

// "<output_dir>/<basename>.block<rank>.<%05d file_number | final>.part"
// (src/output/restart.cpp:54-72)
std::string restart_part_name(std::string const& output_dir, std::string const& basename,
                              int rank, int file_number, bool final_write);

// Combine the parts of one output into "<output_dir>/<basename>.<number|final>
// .restart" (src/output/combine_restart.cpp:77-170, one process): the single
// part is renamed, several are bundled and then removed.  Returns the name of
// the restart file.
std::string combine_restart_parts(std::string const& output_dir, std::string const& basename,
                                  int file_number, bool final_write);

}  // namespace snap
