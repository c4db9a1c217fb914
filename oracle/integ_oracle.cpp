// ORACLE -- test infrastructure only (integration path); filled in below.
extern "C" int oracle_integ_version(void) { return 0; }
