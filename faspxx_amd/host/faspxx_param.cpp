// faspxx_param.cpp -- the Parameters class of the drop-in layer: program
// defaults overridden by a parameter file, overridden by the command line
// (interface src/Param.hxx:82-226; behaviour src/Param.cxx:20-385).
//
//   * Command line: arguments after argv[0] are read as (name, value) pairs;
//     "-h" / "--help" prints the help and exits with PRINT_HELP, an unpaired
//     trailing argument prints it and exits too.  A name given twice keeps
//     its first value.
//   * Parameter file: the first parameter registered with marker 2 names it
//     (its command-line value if one was given, else its current value).  Per
//     line: leading and trailing blanks are dropped; empty lines and lines
//     starting with '#' or '%' are skipped; the first two words are the name
//     and the value, anything after them is ignored.  A line with a name and
//     no value reuses the value of the line before it, and a name given twice
//     keeps its first value, both as the reference's reader behaves.
//   * Parse() records the program's values, then applies the file's values
//     and then the command line's to every registered parameter.  Booleans
//     accept true / false in any case or an integer; unsigned parameters are
//     read as int, as the reference stores them.
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "faspxx_dropin.hxx"

namespace {

// Split one parameter-file line; name and value keep their previous contents
// when the line has no such word.
bool split_line(std::string& line, std::string& name, std::string& value) {
  const size_t first = line.find_first_not_of(' ');
  line.erase(0, first == std::string::npos ? line.size() : first);
  const size_t last = line.find_last_not_of(' ');
  line.erase(last == std::string::npos ? 0 : last + 1);
  if (line.empty() || line[0] == '#' || line[0] == '%') return false;
  std::istringstream words(line);
  words >> name;
  words >> value;
  return true;
}

bool to_bool(const std::string& text) {
  std::string low(text);
  for (char& c : low) c = (char)std::tolower((unsigned char)c);
  if (low == "true") return true;
  if (low == "false") return false;
  return std::stoi(low) != 0;
}

size_t widest(const std::vector<std::string>& names) {
  size_t w = 0;
  for (const auto& n : names) w = std::max(w, n.size());
  return w;
}

}  // namespace

void Parameters::Add(Kind kind, const std::string& name, const std::string& help, void* ptr,
                     int marker) {
  entries.push_back(Entry{kind, name, help, ptr, marker});
}

void Parameters::AddParam(const std::string& name, const std::string& help, bool* ptr,
                          int marker) {
  Add(KBool, name, help, ptr, marker);
}
void Parameters::AddParam(const std::string& name, const std::string& help, USI* ptr,
                          int marker) {
  Add(KInt, name, help, ptr, marker);
}
void Parameters::AddParam(const std::string& name, const std::string& help, INT* ptr,
                          int marker) {
  Add(KInt, name, help, ptr, marker);
}
void Parameters::AddParam(const std::string& name, const std::string& help, double* ptr,
                          int marker) {
  Add(KDouble, name, help, ptr, marker);
}
void Parameters::AddParam(const std::string& name, const std::string& help, std::string* ptr,
                          int marker) {
  Add(KString, name, help, ptr, marker);
}
void Parameters::AddParam(const std::string& name, const std::string& help, Output* ptr,
                          int marker) {
  Add(KOutput, name, help, ptr, marker);
}

std::string Parameters::ValueText(const Entry& e) {
  std::ostringstream s;
  switch (e.kind) {
    case KBool: s << (*(bool*)e.ptr ? "true" : "false"); break;
    case KInt: s << *(int*)e.ptr; break;
    case KDouble: s << *(double*)e.ptr; break;
    case KString: s << *(std::string*)e.ptr; break;
    case KOutput: s << (int)*(Output*)e.ptr; break;
  }
  return s.str();
}

void Parameters::Assign(Entry& e, const std::string& text) {
  switch (e.kind) {
    case KBool: *(bool*)e.ptr = to_bool(text); break;
    case KInt: *(int*)e.ptr = std::stoi(text); break;
    case KDouble: *(double*)e.ptr = std::stod(text); break;
    case KString: *(std::string*)e.ptr = text; break;
    case KOutput: *(Output*)e.ptr = (Output)std::stoi(text); break;
  }
}

void Parameters::ReadCommandLine() {
  for (int i = 1; i < argc; i += 2) {
    if (std::strcmp(argv[i], "-h") == 0 || std::strcmp(argv[i], "--help") == 0) {
      PrintHelp();
      std::exit(PRINT_HELP);
    }
    if (i + 1 >= argc) {
      std::cout << "Check the argument: " << argv[i] << std::endl;
      PrintHelp();
      std::exit(PRINT_HELP);
    }
    fromCommandLine.emplace(argv[i], argv[i + 1]);
  }
}

void Parameters::ReadFile() {
  const auto it = std::find_if(entries.begin(), entries.end(),
                               [](const Entry& e) { return e.marker == 2; });
  if (it == entries.end()) return;
  const auto cl = fromCommandLine.find(it->name);
  const std::string path = cl != fromCommandLine.end() ? cl->second : *(std::string*)it->ptr;
  std::ifstream file(path);
  if (!file.is_open()) std::cout << "### WARNING: File not found!" << std::endl;
  std::string line, name, value;
  while (std::getline(file, line))
    if (split_line(line, name, value)) fromFile.emplace(name, value);
}

void Parameters::Parse() {
  ReadCommandLine();
  ReadFile();
  std::ostringstream s;
  std::vector<std::string> names;
  for (const auto& e : entries) names.push_back(e.name);
  const size_t w = widest(names);
  s << "Original parameters in user program:\n"
    << "---------------------------------------------\n";
  for (const auto& e : entries)
    s << "   " << std::setw((int)w) << std::left << e.name << " [" << ValueText(e) << "]\n";
  s << '\n';
  original = s.str();
  for (auto& e : entries) {
    const auto f = fromFile.find(e.name);
    if (f != fromFile.end()) Assign(e, f->second);
    const auto c = fromCommandLine.find(e.name);
    if (c != fromCommandLine.end()) Assign(e, c->second);
  }
}

void Parameters::SetSOLParams(SOLParams& p) {
  AddParam("-algName", "Name of solver", &p.algName);
  AddParam("-maxIter", "Max iteration steps", &p.maxIter);
  AddParam("-minIter", "Min iteration steps", &p.minIter);
  AddParam("-savIter", "Safe-guard steps", &p.savIter);
  AddParam("-restart", "Restart number", &p.restart);
  AddParam("-resRel", "Relative residual tolerance", &p.relTol);
  AddParam("-resAbs", "Absolute residual tolerance", &p.absTol);
  AddParam("-verbose", "Verbose level", &p.verbose);
}

void Parameters::SetSOLParams(SOLParams& p, SOLParams& pcd) {
  SetSOLParams(p);
  AddParam("-pcdIter", "Preconditioner steps", &pcd.maxIter);
}

void Parameters::PrintUserParams(std::ostream& out) const { out << original << std::endl; }

void Parameters::PrintFileParams(std::ostream& out) const {
  std::vector<std::string> names;
  for (const auto& kv : fromFile) names.push_back(kv.first);
  const size_t w = widest(names);
  std::string path;
  for (const auto& e : entries)
    if (e.marker == 2) path = *(std::string*)e.ptr;
  out << "---------------------------------------------\n"
      << "Parameters from file " << path << ":\n"
      << "---------------------------------------------\n";
  for (const auto& kv : fromFile)
    out << "   " << std::setw((int)w) << std::left << kv.first << " [" << kv.second << "]\n";
  out << std::endl;
}

void Parameters::PrintCommandLineParams(std::ostream& out) const {
  std::vector<std::string> names;
  for (const auto& kv : fromCommandLine) names.push_back(kv.first);
  const size_t w = widest(names);
  out << "---------------------------------------------\n"
      << "Parameters from the command line input\n"
      << "---------------------------------------------\n";
  for (const auto& kv : fromCommandLine)
    out << "   " << std::setw((int)w) << std::left << kv.first << " [" << kv.second << "]\n";
  out << std::endl;
}

void Parameters::Print(std::ostream& out) const {
  std::vector<std::string> names;
  for (const auto& e : entries) names.push_back(e.name);
  const size_t w = widest(names);
  out << "---------------------------------------------\n"
      << "Parameters used in the program\n"
      << "---------------------------------------------\n";
  for (const auto& e : entries)
    out << "   " << std::setw((int)w) << std::left << e.name << " [" << ValueText(e) << "]\n";
  out << '\n';
}

void Parameters::PrintHelp(std::ostream& out) const {
  static const char* kinds[] = {"<bool>", "<int>", "<double>", "<string>", "<Output>"};
  std::vector<std::string> names;
  for (const auto& e : entries) names.push_back(e.name);
  const size_t w = widest(names);
  out << "Usage: " << (argc > 0 ? argv[0] : "program") << " [options] ...\n"
      << "Options:\n";
  for (const auto& e : entries) {
    out << "   " << std::setw((int)w) << std::left << e.name << " " << std::setw(8)
        << kinds[e.kind];
    out << (e.marker == 0 ? ", optional  " : e.marker == 1 ? ", required  " : ", from file ");
    if (!e.help.empty()) out << " :  " << e.help;
    out << ", default = [" << ValueText(e) << "]\n";
  }
}
