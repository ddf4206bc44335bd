// placeholder
int main() { return 0; }
