// Restart files: TorchScript payloads, the bundle index, and the naming and
// combining rules of the reference's restart output.
#include "restart.hpp"

#include <torch/script.h>

#include <algorithm>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace fs = std::filesystem;

namespace snap {
namespace {

constexpr char kBundleMagic[] = "SNAPY_RESTART_BUNDLE_V1";

bool has_suffix(std::string const& s, std::string const& suffix) {
  return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

bool has_prefix(std::string const& s, std::string const& prefix) {
  return s.size() >= prefix.size() && s.compare(0, prefix.size(), prefix) == 0;
}

Variables buffers_of(torch::jit::Module const& m) {
  Variables out;
  for (auto const& b : m.named_buffers()) out[b.name] = b.value;
  return out;
}

// block rank named by "<basename>.block<rank>.<number>.part"
// (src/input/read_restart_file.cpp:36-65, 172-173); malformed names throw
int part_block_rank(std::string const& name) {
  if (name.size() <= 5 || !has_suffix(name, ".part"))
    throw std::invalid_argument("filename does not end with .part");
  const std::string core = name.substr(0, name.size() - 5);
  const size_t dot2 = core.rfind('.');
  if (dot2 == std::string::npos) throw std::invalid_argument("filename missing filenumber field");
  const size_t dot1 = dot2 == 0 ? std::string::npos : core.rfind('.', dot2 - 1);
  if (dot1 == std::string::npos) throw std::invalid_argument("filename missing block_id field");
  const std::string block = core.substr(dot1 + 1, dot2 - dot1 - 1);
  if (dot1 == 0 || block.empty() || dot2 + 1 == core.size())
    throw std::invalid_argument("one or more filename fields are empty");
  if (!has_prefix(block, "block") || block.size() == 5 ||
      block.find_first_not_of("0123456789", 5) != std::string::npos)
    throw std::invalid_argument("block_id field '" + block + "' is not block<rank>");
  return std::stoi(block.substr(5));
}

struct BundleEntry {
  std::string name;
  uint64_t size = 0;
  uint64_t offset = 0;
};

// index of a bundle whose magic line has been read
std::vector<BundleEntry> read_bundle_index(std::istream& in) {
  std::string line;
  if (!std::getline(in, line) || line.empty())
    throw std::invalid_argument("restart bundle missing entry count");
  const uint64_t count = std::stoull(line);
  std::vector<BundleEntry> entries;
  for (uint64_t i = 0; i < count; ++i) {
    if (!std::getline(in, line)) throw std::invalid_argument("restart bundle index is truncated");
    const size_t tab = line.find('\t');
    if (tab == std::string::npos) throw std::invalid_argument("restart bundle index line missing tab");
    BundleEntry e;
    e.name = line.substr(0, tab);
    e.size = std::stoull(line.substr(tab + 1));
    entries.push_back(std::move(e));
  }
  if (!std::getline(in, line) || !line.empty())
    throw std::invalid_argument("restart bundle header missing terminator");
  uint64_t offset = static_cast<uint64_t>(in.tellg());
  for (auto& e : entries) {
    e.offset = offset;
    offset += e.size;
  }
  return entries;
}

Variables load_from_bundle(std::ifstream& in, std::string const& path, int block_rank) {
  std::vector<BundleEntry> entries;
  try {
    entries = read_bundle_index(in);
  } catch (std::exception const& exc) {
    std::cerr << path << ": failed to parse restart bundle: " << exc.what() << "\n";
    return {};
  }
  for (auto const& e : entries) {
    if (!has_suffix(e.name, ".part")) continue;
    if (part_block_rank(e.name) != block_rank) continue;
    // the payload is read into memory and loaded from there (no temporary file)
    std::string payload(e.size, '\0');
    in.clear();
    in.seekg(static_cast<std::streamoff>(e.offset));
    in.read(payload.data(), static_cast<std::streamsize>(e.size));
    if (!in || static_cast<uint64_t>(in.gcount()) != e.size) {
      std::cerr << path << ": truncated restart bundle payload\n";
      return {};
    }
    std::istringstream ss(payload, std::ios::binary);
    return buffers_of(torch::jit::load(ss, torch::kCPU));
  }
  std::cerr << path << ": no matching block found in restart bundle\n";
  return {};
}

}  // namespace

void save_tensors(Variables const& vars, std::string const& filename) {
  torch::jit::Module m("TensorModule");
  for (auto const& [name, t] : vars)
    if (t.defined()) m.register_buffer(name, t);
  m.save(filename);
}

Variables load_tensors(std::string const& filename) {
  return buffers_of(torch::jit::load(filename, torch::kCPU));
}

int make_restart_bundle(std::string const& bundle_name, std::vector<std::string> const& file_list) {
  std::vector<uint64_t> sizes;
  for (auto const& f : file_list) {
    std::error_code ec;
    auto n = fs::file_size(f, ec);
    if (ec) return -1;
    sizes.push_back(n);
  }
  std::ofstream out(bundle_name, std::ios::binary | std::ios::trunc);
  if (!out) return -1;
  out << kBundleMagic << "\n" << file_list.size() << "\n";
  for (size_t i = 0; i < file_list.size(); ++i)
    out << fs::path(file_list[i]).filename().string() << "\t" << sizes[i] << "\n";
  out << "\n";
  std::vector<char> buf(1 << 20);
  for (size_t i = 0; i < file_list.size(); ++i) {
    std::ifstream in(file_list[i], std::ios::binary);
    if (!in) return -1;
    uint64_t copied = 0;
    while (in) {
      in.read(buf.data(), static_cast<std::streamsize>(buf.size()));
      const auto n = in.gcount();
      if (n > 0) out.write(buf.data(), n), copied += static_cast<uint64_t>(n);
    }
    // the index promised sizes[i] bytes: a file that changed size meanwhile
    // would shift every later payload
    if (copied != sizes[i]) return -1;
  }
  out.flush();
  return out ? 0 : -1;
}

Variables load_restart(std::string const& path, int block_rank) {
  {
    // a bundle starts with the magic line; read only that many bytes (a
    // payload is a zip archive and may have no newline for megabytes)
    std::ifstream in(path, std::ios::binary);
    const std::string magic = std::string(kBundleMagic) + "\n";
    std::string head(magic.size(), '\0');
    if (in && in.read(&head[0], (std::streamsize)head.size()) && head == magic)
      return load_from_bundle(in, path, block_rank);
  }
  return load_tensors(path);
}

std::string restart_part_name(std::string const& output_dir, std::string const& basename,
                              int rank, int file_number, bool final_write) {
  char number[32];
  std::snprintf(number, sizeof(number), "%05d", file_number);
  return output_dir + "/" + basename + ".block" + std::to_string(rank) + "." +
         (final_write ? std::string("final") : std::string(number)) + ".part";
}

std::string combine_restart_parts(std::string const& output_dir, std::string const& basename,
                                  int file_number, bool final_write) {
  char number[32];
  std::snprintf(number, sizeof(number), "%05d", file_number);
  const std::string tag = final_write ? std::string("final") : std::string(number);
  const std::string prefix = basename + ".block", suffix = "." + tag + ".part";
  std::vector<std::string> parts;
  std::error_code ec;
  for (auto const& entry : fs::directory_iterator(output_dir, ec)) {
    const std::string name = entry.path().filename().string();
    if (name.size() > prefix.size() + suffix.size() && has_prefix(name, prefix) &&
        has_suffix(name, suffix))
      parts.push_back(entry.path().string());
  }
  if (ec || parts.empty())
    throw std::runtime_error("combine_restart_parts: no '" + prefix + "*" + suffix + "' in '" +
                             output_dir + "'");
  std::sort(parts.begin(), parts.end());
  const std::string outfile = output_dir + "/" + basename + "." + tag + ".restart";
  std::remove(outfile.c_str());
  int err = parts.size() == 1 ? std::rename(parts.front().c_str(), outfile.c_str())
                              : make_restart_bundle(outfile, parts);
  if (err) throw std::runtime_error("restart bundling failed with error " + std::to_string(err));
  if (parts.size() != 1)
    for (auto const& p : parts) std::remove(p.c_str());
  return outfile;
}

}  // namespace snap
