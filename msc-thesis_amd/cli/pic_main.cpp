// placeholder, replaced by the CLI
int main() { return 0; }
