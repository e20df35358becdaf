/* Oracle harness entry point (test infrastructure only): starts the Fortran
 * runtime with the process arguments and hands control to the Fortran driver in
 * ref_driver.f90, which runs the compiled reference code. */
void _FortranAProgramStart(int argc, const char *argv[], const char *envp[], const void *env);
void _FortranAProgramEndStatement(void);
void ref_driver_main(void);
int main(int argc, const char *argv[], const char *envp[]) {
  _FortranAProgramStart(argc, argv, envp, 0);
  ref_driver_main();
  _FortranAProgramEndStatement();
  return 0;
}
