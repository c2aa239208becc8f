"""MI355X-native engine for the likelihood of codeml's branch-site model A."""
