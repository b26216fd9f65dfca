// param_tests.cpp -- the reference's Parameters unit test
// (utest/src/UnitTestsParam.cxx:17-93) on the drop-in class, plus the parameter
// file rules that test leaves open.  Host code only; no device is touched.
//   param_tests <golden dir>
// Exit status: the number of failed checks.
#include <cmath>
#include <cstdio>
#include <string>

#include "Param.hxx"

static int failures = 0;
#define CHECK(cond)                                                      \
  do {                                                                   \
    if (!(cond)) {                                                       \
      std::printf("FAILED line %d: %s\n", __LINE__, #cond);              \
      ++failures;                                                        \
    }                                                                    \
  } while (0)

// UnitTestsParam.cxx:17-93: command line over file over program values.
static void unit_test(const std::string& dir) {
  bool bool_param = false;
  INT int_param = 11;
  DBL double_param = 3.14159;
  std::string char_param = "user params";
  Output output_lvl = static_cast<Output>(4);
  const std::string single = dir + "/single_sol.param";
  const char* argv[] = {"program name", "-bool_param", "TruE", "-int_param", "22",
                        "-double_param", "1.41414", "-char_param", "commandline_parameters",
                        "-faspxx_opts", single.c_str()};
  std::string params_file = dir + "/multiple_sol.param";
  bool view_param = false;
  INT level_param = 0;
  DBL resrel_param = 0.0;
  std::string vec_param = "";

  Parameters params(11, argv);
  params.AddParam("-bool_param", "bool param help", &bool_param);
  params.AddParam("-int_param", "integer param help", &int_param);
  params.AddParam("-double_param", "double param help", &double_param);
  params.AddParam("-char_param", "char* param help", &char_param);
  params.AddParam("-faspxx_opts", "additional parameter file", &params_file, 2);
  params.AddParam("-view", "view from file", &view_param);
  params.AddParam("-level", "level from file", &level_param);
  params.AddParam("-resrel", "resrel from file", &resrel_param);
  params.AddParam("-vec", "vec from file", &vec_param);
  params.AddParam("-out_put", "output param help", &output_lvl);
  params.Parse();

  CHECK(bool_param == true);
  CHECK(int_param == 22);
  CHECK(double_param == 1.41414);
  CHECK(char_param == "commandline_parameters");
  CHECK(params_file == single);
  CHECK(view_param == true);
  CHECK(level_param == 4);
  CHECK(std::abs(resrel_param - 1.234e-6) < 1E-14);
  CHECK(vec_param == "../data/ffffffffffff");
  CHECK(output_lvl == 6);
}

// The file named by the marker-2 parameter's own value (no command-line
// override); a name without a value takes the previous line's value; words
// after the value are ignored; unsigned and Output parameters; a file name
// given twice keeps its first value.
static void file_rules(const std::string& dir) {
  std::string file = dir + "/multiple_sol.param";
  std::string type = "none", solver2_type = "none";
  DBL rtol = 0.0;
  USI maxiter = 0, restart2 = 0;
  INT view = -1, solver2_view = -1;
  Output printlvl = PRINT_MAX;
  const char* argv[] = {"prog", "-sol_maxiter", "17"};
  Parameters p(3, argv);
  p.AddParam("-opts", "parameter file", &file, 2);
  p.AddParam("-sol_type", "", &type);
  p.AddParam("-sol_rtol", "", &rtol);
  p.AddParam("-sol_maxiter", "", &maxiter);
  p.AddParam("-sol_printlvl", "", &printlvl);
  p.AddParam("-sol_view", "", &view);
  p.AddParam("-solver2_sol_type", "", &solver2_type);
  p.AddParam("-solver2_sol_restart", "", &restart2);
  p.AddParam("-solver2_sol_view", "", &solver2_view);
  p.Parse();
  CHECK(type == "cg");           // "cg       # ..." : the comment is ignored
  CHECK(rtol == 1.23E-8);        // "  -sol_rtol   1.23E-8": leading blanks dropped
  CHECK(maxiter == 17);          // the command line wins over the file's 789
  CHECK(printlvl == PRINT_NONE + 1);
  CHECK(view == 1);              // "-sol_view" alone: the previous line's "1"
  CHECK(solver2_type == "gmres");
  CHECK(restart2 == 30);
  CHECK(solver2_view == 0);      // after "-solver2_sol_printlvl 0"
}

// SetSOLParams registers the solver options under the reference's names.
static void sol_params() {
  SOLParams sp, pp;
  const char* argv[] = {"prog", "-maxIter", "150", "-resRel", "1e-5", "-algName", "gmres",
                        "-pcdIter", "3", "-verbose", "8", "-restart", "25"};
  Parameters p(13, argv);
  p.SetSOLParams(sp, pp);
  p.Parse();
  CHECK(sp.maxIter == 150);
  CHECK(sp.relTol == 1e-5);
  CHECK(sp.absTol == 1e-8);  // default kept
  CHECK(sp.algName == "gmres");
  CHECK(sp.restart == 25);
  CHECK(sp.verbose == PRINT_MAX);
  CHECK(pp.maxIter == 3);
}

int main(int argc, char** argv) {
  if (argc != 2) return 100;
  const std::string dir = argv[1];
  unit_test(dir);
  file_rules(dir);
  sol_params();
  std::printf("%d failures\n", failures);
  return failures;
}
