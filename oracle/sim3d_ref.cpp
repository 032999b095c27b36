// ORACLE — 3D restatement (filled in later)
